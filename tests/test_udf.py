"""Wall values given as expressions (UDF_D/U/V/T, geom_bc_defs.py:115-146).

The host compiles each string to a stack program for the device; the oracle
walks the same string as a C expression tree with numpy.  Here the compiled
program, run by a Python restatement of the device interpreter, must give the
oracle's values (C integer semantics included); malformed programs must be
refused by ugks_block_create before anything reaches the device; and the
oracle's paraBC with expressions is pinned by the constant case and by the
closed-form wall state of a time-dependent lid.
"""
import copy
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from pyugks_amd import bc as bcd
from pyugks_amd import cases, udf
from pyugks_amd import lib as L

EXPRS = ["1.0", "0.15", "1/2", "1/2 + x", "7 % 3 * t", "-7/2 + s", "2*PI*t", "0.1*sin(2*PI*t/0.05) + 0.05*s",
         "1.0 + 0.1*x - 0.2*y*y", "pow(x, 2) + fmin(y, t) - -3", "sqrt(x*x + y*y)", "exp(-t)*cos(3*s)",
         "fmax(0.5, fabs(x - 0.5))", "atan2(y, x + 1.0)", "fmod(t, 0.3) + floor(4*s) / 4", "tanh(10*(x - 0.5))",
         "log(1 + s) + log10(2 + t)", "(1 + 2) * (3 - t) / (4 + s)", "+x - -y", "ceil(2.5*s) + asin(0.5*s)",
         "acos(0.3*y) + sinh(t) - cosh(0.1*x) + tan(0.2*s) + atan(x)"]


def run_program(code, consts, s, x, y, t):
    """The device interpreter (udf_eval) restated in Python, for the compiler's tests."""
    st = []
    i = 0
    un = {14: lambda a: -a, 20: np.sin, 21: np.cos, 22: np.tan, 23: np.exp, 24: np.log, 25: np.sqrt, 26: np.abs,
          27: np.tanh, 28: np.arctan, 29: np.arcsin, 30: np.arccos, 31: np.sinh, 32: np.cosh, 33: np.log10,
          34: np.floor, 35: np.ceil}
    bi = {10: np.add, 11: np.subtract, 12: np.multiply, 13: np.divide, 15: np.power, 40: np.fmin, 41: np.fmax,
          42: np.arctan2, 43: np.fmod}
    while i < len(code):
        op = code[i]
        if op == udf.OP_CONST:
            st.append(np.float64(consts[code[i + 1]]))
            i += 2
            continue
        if op in (2, 3, 4, 5):
            st.append({2: s, 3: x, 4: y, 5: t}[op])
        elif op in bi:
            b, a = st.pop(), st.pop()
            st.append(bi[op](a, b))
        else:
            st.append(un[op](st.pop()))
        i += 1
    assert len(st) == 1
    return st[0]


@pytest.mark.parametrize("text", EXPRS)
def test_compiled_program_matches_the_c_expression(text):
    rng = np.random.default_rng(1)
    s, x, y = rng.random(16), rng.random(16), rng.random(16)
    t = 0.37
    code, consts = udf.compile_udf(text)
    got = run_program(code, consts, s, x, y, np.float64(t))
    want = O.udf_values(text, s, x, y, t)
    assert np.allclose(got, want, rtol=1e-14, atol=1e-15), text


def test_c_integer_semantics():
    assert udf.compile_udf("1/2")[1] == [0.0]
    assert udf.compile_udf("-7/2")[1] == [-3.0]
    assert udf.compile_udf("7 % 3")[1] == [1.0]
    assert udf.compile_udf("-7 % 3")[1] == [-1.0]
    assert udf.compile_udf("1.0/2")[1] == [0.5]
    assert O.udf_values("1/2 + 0*s", np.zeros(2), 0, 0, 0)[0] == 0.0
    assert O.udf_values("-7 % 3 + 0*s", np.zeros(1), 0, 0, 0)[0] == -1.0


@pytest.mark.parametrize("text", ["z + 1", "x ** 2", "foo(x)", "sin(x, y)", "x % 2", "x if t else y", "1/0", "("])
def test_bad_expressions_are_refused(text):
    with pytest.raises(udf.UDFError):
        udf.compile_udf(text)


def test_deep_expression_is_refused():
    text = "x"
    for _ in range(20):
        text = f"(1 + {text} * (2 + x))"
    with pytest.raises(udf.UDFError):
        udf.compile_udf(text)


def _desc_with_program(code, consts, slot=2, length=None):
    g = cases.cavity(4, ("NC", 8), "vanLeer", 0.05, n_cells_j=4).prepare()
    c = g.block_constants(g.blocks[0])
    code = np.ascontiguousarray(code, dtype=np.int32)
    consts = np.ascontiguousarray(consts or [0.0], dtype=np.float64)
    start, lens = [0] * 16, [0] * 16
    lens[slot] = len(code) if length is None else length
    c["udf"] = (code, consts, start, lens)
    return L.make_desc(c, g.velocity_tables())


@pytest.mark.parametrize("code,consts,why", [
    ([1, 0, 99], [1.0], "unknown opcode"),
    ([1, 3], [1.0], "constant index"),
    ([10], [], "underflow"),
    ([3, 4], [], "exactly one value"),
    ([3] * 17 + [10] * 16, [], "deeper than 16"),
])
def test_block_create_refuses_malformed_programs(code, consts, why):
    """The checks run before any device call, so this needs no GPU."""
    desc, keep = _desc_with_program(code, consts)
    h = C.c_void_p()
    rc = L.load().ugks_block_create(C.byref(desc), C.byref(h))
    assert rc != 0
    assert why in L.load().ugks_last_error().decode()


def test_block_create_refuses_a_program_outside_its_table():
    desc, keep = _desc_with_program([3], [], length=4)
    h = C.c_void_p()
    assert L.load().ugks_block_create(C.byref(desc), C.byref(h)) != 0
    assert "code table" in L.load().ugks_last_error().decode()


# --------------------------------------------------- oracle paraBC pins --
def _udf_cavity(**lid):
    g = cases.cavity(8, ("NC", 8), "WENO5", 0.05, lid=0.15, n_cells_j=6)
    g.blocks[0].bc_list[bcd.NORTH] = bcd.DiffuseBC(D=1.0, T=1.0, U=0.15, **lid)
    return g


def test_constant_expressions_run_as_constants():
    """UDF strings that are constants give the constant-wall run bit for bit."""
    a = O.OracleSim(_udf_cavity(), threads=2)
    b = O.OracleSim(_udf_cavity(UDF_U="0.15", UDF_T="1.0", UDF_D="2/2", UDF_V="0*t"), threads=2)
    assert b.time_dependent == [True]
    a.run(3)
    b.run(3)
    assert np.array_equal(a.blocks[0]["f"], b.blocks[0]["f"])
    assert np.array_equal(a.blocks[0]["wall_prop"], b.blocks[0]["wall_prop"])


def test_time_dependent_lid_takes_its_closed_form_value():
    """A lid U = 0.2 sin(2 pi t / 0.05) + 0.05 s with T = 1 + 0.1 x: after each
    step the north wall state holds those values at the step's start time,
    at each lid cell's arc position s = (i + 1/2)/n and midpoint x."""
    o = O.OracleSim(_udf_cavity(UDF_U="0.2*sin(2*PI*t/0.05) + 0.05*s", UDF_T="1.0 + 0.1*x"), threads=2)
    n = 8
    s = (np.arange(n) + 0.5) / n
    for _ in range(3):
        t0 = o.time
        o.one_step(True)
        wp = o.blocks[0]["wall_prop"][:n, bcd.NORTH]
        assert np.allclose(wp[:, 1], 0.2 * np.sin(2 * math.pi * t0 / 0.05) + 0.05 * s, rtol=0, atol=1e-15)
        assert np.allclose(wp[:, 3], 1.0 / (1.0 + 0.1 * s), rtol=1e-15)
        assert np.all(wp[:, 0] == 1.0) and np.all(wp[:, 2] == 0.0)
    # the lid really moves the gas differently from a constant lid
    c = O.OracleSim(_udf_cavity(), threads=2)
    c.run(3)
    assert np.max(np.abs(c.blocks[0]["macro"] - o.blocks[0]["macro"])) > 1e-6
