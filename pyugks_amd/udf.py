"""Wall values given as expressions: the UDF_D / UDF_U / UDF_V / UDF_T strings
of a boundary condition (geom/geom_bc_defs.py:115-146).

The reference pastes each string into the kernel source as a #define
(ugks_CL.py:278-298), so it is an OpenCL C expression in the wall cell's
parametric position s (0..1 along the wall), its midpoint (x, y) and the
simulation time t, with the OpenCL math functions and the PI macro; paraBC
evaluates it once per wall cell (clKernel.c:141-218).  Here the string is
compiled once, on the host, to a short stack program that the device
evaluates in the same place (no run-time compiler): C semantics are kept
where they differ from Python's -- integer literals stay integers, so
``1/2`` is 0 and ``7 % 3`` is 1, and mixed operands promote to double.

Opcodes (int32; CONST is followed by an index into the constant table):
"""
from __future__ import annotations

import ast
import math

import numpy as np

OP_CONST, OP_S, OP_X, OP_Y, OP_T = 1, 2, 3, 4, 5
OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_NEG, OP_POW = 10, 11, 12, 13, 14, 15
UNARY = {"sin": 20, "cos": 21, "tan": 22, "exp": 23, "log": 24, "sqrt": 25, "fabs": 26, "tanh": 27, "atan": 28,
         "asin": 29, "acos": 30, "sinh": 31, "cosh": 32, "log10": 33, "floor": 34, "ceil": 35}
BINARY = {"pow": OP_POW, "fmin": 40, "fmax": 41, "atan2": 42, "fmod": 43}
MAX_STACK = 16
MAX_OPS = 256
PI_BAKED = 3.141592653589793     # '%0.15e' % pi, the PI macro of ugks_CL.py:84
VARS = {"s": OP_S, "x": OP_X, "y": OP_Y, "t": OP_T}

__doc__ += "".join(f"  {k:>6} {v}\n" for k, v in
                   [("CONST", OP_CONST), ("s", OP_S), ("x", OP_X), ("y", OP_Y), ("t", OP_T), ("+", OP_ADD),
                    ("-", OP_SUB), ("*", OP_MUL), ("/", OP_DIV), ("neg", OP_NEG), ("pow", OP_POW)]
                   + list(UNARY.items()) + [(k, v) for k, v in BINARY.items() if k != "pow"])


class UDFError(ValueError):
    pass


def _c_int_div(a: int, b: int) -> int:
    if b == 0:
        raise UDFError("integer division by zero in a wall-value expression")
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def _c_int_mod(a: int, b: int) -> int:
    return a - _c_int_div(a, b) * b


class _Compiler:
    """Constant subexpressions are folded with C rules; the rest becomes a
    post-order list of opcodes.  A node compiles to ("const", value, is_int)
    or ("code", words, stack depth)."""

    def __init__(self, consts=None):
        self.consts: list[float] = [] if consts is None else consts

    def node(self, n):
        if isinstance(n, ast.Expression):
            return self.node(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)) and not isinstance(n.value, bool):
            return ("const", n.value, isinstance(n.value, int))
        if isinstance(n, ast.Name):
            if n.id in VARS:
                return ("code", [VARS[n.id]], 1)
            if n.id in ("PI", "M_PI"):
                return ("const", PI_BAKED, False)
            raise UDFError(f"unknown name {n.id!r} in a wall-value expression (s, x, y, t and PI are known)")
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd)):
            a = self.node(n.operand)
            if a[0] == "const":
                return ("const", -a[1] if isinstance(n.op, ast.USub) else a[1], a[2])
            if isinstance(n.op, ast.UAdd):
                return a
            return ("code", a[1] + [OP_NEG], a[2])
        if isinstance(n, ast.BinOp):
            ops = {ast.Add: OP_ADD, ast.Sub: OP_SUB, ast.Mult: OP_MUL, ast.Div: OP_DIV, ast.Mod: 43}
            k = ops.get(type(n.op))
            if k is None:
                raise UDFError(f"operator {type(n.op).__name__} is not an OpenCL C arithmetic operator here")
            a, b = self.node(n.left), self.node(n.right)
            if a[0] == "const" and b[0] == "const":
                return self.fold(k, a, b)
            if k == 43:
                raise UDFError("% needs integer operands in C; use fmod() for doubles")
            return self.join([a, b], k)
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and not n.keywords:
            f = n.func.id
            args = [self.node(x) for x in n.args]
            if f in UNARY and len(args) == 1:
                if args[0][0] == "const":
                    return ("const", _host_unary(f, float(args[0][1])), False)
                return self.join(args, UNARY[f])
            if f in BINARY and len(args) == 2:
                if args[0][0] == "const" and args[1][0] == "const":
                    return ("const", _host_binary(f, float(args[0][1]), float(args[1][1])), False)
                return self.join(args, BINARY[f])
            raise UDFError(f"function {f}() with {len(args)} argument(s) is not supported in a wall-value expression")
        raise UDFError(f"unsupported syntax in a wall-value expression: {ast.dump(n)[:60]}")

    def fold(self, k, a, b):
        x, y = a[1], b[1]
        if a[2] and b[2]:            # int op int stays int (C)
            if k == OP_ADD:
                return ("const", x + y, True)
            if k == OP_SUB:
                return ("const", x - y, True)
            if k == OP_MUL:
                return ("const", x * y, True)
            if k == OP_DIV:
                return ("const", _c_int_div(x, y), True)
            return ("const", _c_int_mod(x, y), True)
        x, y = float(x), float(y)
        if k == 43:
            raise UDFError("% needs integer operands in C; use fmod() for doubles")
        v = {OP_ADD: lambda: x + y, OP_SUB: lambda: x - y, OP_MUL: lambda: x * y,
             OP_DIV: lambda: x / y if y != 0.0 else math.copysign(math.inf, x) * math.copysign(1.0, y)}[k]()
        return ("const", v, False)

    def words(self, a):
        """(opcodes, stack depth) of a compiled node, constants placed in the table."""
        if a[0] == "const":
            self.consts.append(float(a[1]))
            return [OP_CONST, len(self.consts) - 1], 1
        return a[1], a[2]

    def join(self, args, op):
        """The operands in order, then op: depth max over operand k of k + its own."""
        out, depth = [], 0
        for k, a in enumerate(args):
            w, d = self.words(a)
            out += w
            depth = max(depth, k + d)
        if depth > MAX_STACK:
            raise UDFError(f"wall-value expression nests deeper than {MAX_STACK}")
        return ("code", out + [op], depth)


def _host_unary(f, x):
    return float(getattr(np, {"fabs": "abs", "asin": "arcsin", "acos": "arccos", "atan": "arctan"}.get(f, f))(x))


def _host_binary(f, x, y):
    return float({"pow": np.power, "fmin": np.fmin, "fmax": np.fmax, "atan2": np.arctan2, "fmod": np.fmod}[f](x, y))


def compile_udf(text: str, consts=None):
    """(code, consts) of one wall-value expression; a constant expression
    compiles to a single CONST.  Constants are appended to ``consts`` when it
    is given (a table shared by several programs)."""
    try:
        tree = ast.parse(text.strip(), mode="eval")
    except SyntaxError as e:
        raise UDFError(f"cannot parse wall-value expression {text!r}: {e.msg}") from None
    c = _Compiler(consts)
    code, _ = c.words(c.node(tree))
    if len(code) > MAX_OPS:
        raise UDFError(f"wall-value expression longer than {MAX_OPS} opcodes")
    return code, c.consts


def is_time_dependent(bc) -> bool:
    """ugks_block.py:114-119: a block re-evaluates its wall values every step
    when any UDF string of its boundaries contains the letter t (a substring
    test, so sqrt() counts too)."""
    return any(u is not None and "t" in u for u in (bc.UDF_D, bc.UDF_T, bc.UDF_U, bc.UDF_V))


def block_programs(bc_list):
    """The UDF programs of a block's four faces: (code, consts, start[16],
    length[16]) with slot 4 * face + q, q = 0 D, 1 U, 2 V, 3 T; length 0 where
    the constant wall value stands."""
    code, consts = [], []
    start, length = [0] * 16, [0] * 16
    for w, bc in enumerate(bc_list):
        if bc is None:
            continue
        for q, name in enumerate(("UDF_D", "UDF_U", "UDF_V", "UDF_T")):
            text = getattr(bc, name, None)
            if text is None:
                continue
            c, _ = compile_udf(text, consts)
            start[4 * w + q], length[4 * w + q] = len(code), len(c)
            code += c
    return (np.ascontiguousarray(code or [0], dtype=np.int32), np.ascontiguousarray(consts or [0.0], dtype=np.float64),
            start, length)


def para_table(ni: int, nj: int, G: int, side_length: np.ndarray, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """The parametric position of each wall cell, [max(ni, nj)][4][4] = (s, x,
    y, 0) per wall (ugks_block.py:776-836): s the arc length to the cell's
    midpoint over the wall's length, (x, y) the midpoint of its two vertices.
    The west wall takes its vertices from the second column of the grid
    (self.x[1, :], ugks_block.py:786), as the reference does."""
    para = np.zeros((max(ni, nj), 4, 4), dtype=np.float64)
    L = np.asarray(side_length, dtype=np.float64)
    walls = {0: (L[G:-G, -G, 0], x[:, -1], y[:, -1]), 2: (L[G:-G, G, 0], x[:, 0], y[:, 0]),
             1: (L[-G, G:-G, 1], x[-1, :], y[-1, :]), 3: (L[G, G:-G, 1], x[1, :], y[1, :])}
    for w, (lens, xs, ys) in walls.items():
        total = np.sum(lens)
        length = 0.0
        for i in range(len(lens)):
            para[i, w, 0] = (length + lens[i] / 2.0) / total
            para[i, w, 1] = (xs[i] + xs[i + 1]) / 2.0
            para[i, w, 2] = (ys[i] + ys[i + 1]) / 2.0
            length += lens[i]
    return para
