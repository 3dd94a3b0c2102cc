"""ctypes driver of the FP64 CPU oracle (oracle/ugks_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke()
and by the cpu_baseline leg of bench.py, as the checker.  The product
(pyugks_amd) never imports it.

``OracleSim(gdata)`` builds the same blocks as ``pyugks_amd.sim.UGKSim``
from the same case definition, holds every buffer as a numpy array in the
reference layout, and runs the reference sequence in C.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
SRC = os.path.join(HERE, "ugks_oracle.c")


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (OpenMP, no FP contraction)."""
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
        tmp = LIB + ".tmp"
        subprocess.run(["gcc", "-O2", "-std=c11", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off",
                        "-o", tmp, SRC, "-lm"], check=True)
        os.replace(tmp, LIB)
    return LIB


P = C.c_void_p
I = C.c_int32
D = C.c_double


class Params(C.Structure):
    _fields_ = [
        ("ni", I), ("nj", I), ("NI", I), ("NJ", I), ("NV", I), ("ghost", I), ("flux_method", I), ("relax_type", I),
        ("Pr", D), ("Kn_eff", D), ("gam", D), ("B", D), ("K", D), ("omega", D), ("CFL", D), ("umax", D), ("vmax", D),
        ("wall_D", D * 4), ("wall_U", D * 4), ("wall_V", D * 4), ("wall_T", D * 4), ("wall_P", D * 4),
        ("beta_n", D * 4), ("beta_t", D * 4), ("gamma_f", D * 4), ("alpha_p", D * 4),
        ("dynamic_vartheta", I * 4), ("constant_adsorb", I * 4), ("adsorb", D * 4),
        ("quad", P), ("weight", P), ("mirror_NS", P), ("mirror_EW", P), ("mirror_D", P),
        ("alpha_n", D * 4), ("alpha_t", D * 4),
        ("iso", P * 4), ("iso_tri", P * 4), ("iso_nbr", P * 4), ("n_iso", I * 4), ("n_tri", I * 4),
    ]


STATE_FIELDS = ["f", "flux_f_S", "flux_f_W", "sigma", "macro", "Q", "flux_macro_S", "flux_macro_W", "prim", "aL",
                "aR", "aT", "faceQ", "Mxi", "xy", "area", "centre", "mid_side", "normal", "side_length", "wall_prop",
                "wall_cover", "wall_fluxes", "wall_dist", "time_step", "residual", "bc_mdot", "flag"]


class State(C.Structure):
    _fields_ = [(n, P) for n in STATE_FIELDS]


class BC(C.Structure):
    _fields_ = [("type", I), ("other_block", I), ("other_face", I), ("orientation", I), ("flip", I),
                ("reflect_type", I)]


class Block(C.Structure):
    _fields_ = [("p", Params), ("s", State), ("bc", BC * 4)]


class Sim(C.Structure):
    _fields_ = [("nblocks", I), ("blocks", C.POINTER(Block)), ("dt", D), ("time", D), ("CFL", D),
                ("delta_dt", D), ("suggest_dt", D), ("Kn_eff", D), ("step", I), ("force_time", I),
                ("dt_update_count", I), ("max_dt", P)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(build())
        _lib.or_sim_init.argtypes = [C.POINTER(Sim)]
        _lib.or_block_init.argtypes = [C.POINTER(Block)]
        _lib.or_sim_updateAllBC.argtypes = [C.POINTER(Sim)]
        _lib.or_sim_calcTimeStep.argtypes = [C.POINTER(Sim)]
        _lib.or_sim_updateTimeStep.argtypes = [C.POINTER(Sim), C.c_int]
        _lib.or_cellGeom.argtypes = [C.c_void_p] * 7
        _lib.or_sim_one_step.argtypes = [C.POINTER(Sim), C.c_int]
        _lib.or_sim_run.argtypes = [C.POINTER(Sim), C.c_int]
        _lib.or_set_threads.argtypes = [C.c_int]
        _lib.or_getInternal.argtypes = [C.c_void_p] * 4
        for n in ("or_sim_init", "or_sim_one_step", "or_sim_run"):
            getattr(_lib, n).restype = C.c_int
    return _lib


def set_threads(n: int):
    lib().or_set_threads(int(n))


class V2(C.Structure):
    _fields_ = [("x", D), ("y", D)]


class V4(C.Structure):
    _fields_ = [("s0", D), ("s1", D), ("s2", D), ("s3", D)]


def helpers():
    """Bind the oracle's scalar helpers (for checks against closed forms)."""
    L = lib()
    PP = C.POINTER(Params)
    sig = {
        "or_relaxTime": (D, [PP, V4]),
        "or_fM": (V2, [PP, V4, V2, C.c_int]),
        "or_fS": (V2, [PP, V4, V2, V2, V2]),
        "or_momentU": (None, [PP, V4, C.POINTER(D), C.POINTER(D), C.POINTER(D), C.POINTER(D), C.POINTER(D)]),
        "or_timeIntegrals": (None, [D, D, C.POINTER(D)]),
        "or_vanLeer": (V2, [V2, V2]),
        "or_WENO5": (V2, [C.POINTER(V2), D]),
        "or_NND": (V2, [C.POINTER(V2), D]),
        "or_microSlope": (V4, [PP, V4, V4]),
        "or_getPrimary": (V4, [PP, V4]),
        "or_getConserved": (V4, [PP, V4]),
        "or_bessel_i0": (D, [D]),
        "or_langmuir_table": (D, [PP, C.c_int, D, D, C.POINTER(C.c_int)]),
        "or_cl_kernel": (D, [V2, V2, D, D, D]),
        "or_adsorbingWallCL_P2": (None, [PP, P, C.c_int, P, P, P, P, P, D]),
        "or_adsorbingWallD_P2": (None, [PP, P, C.c_int, P, P, P, P, P, D]),
    }
    for n, (r, a) in sig.items():
        f = getattr(L, n)
        f.restype = r
        f.argtypes = a
    return L


def make_params(consts: dict, tables) -> Params:
    q, w, mns, mew, md = tables
    p = Params()
    for k in ("ni", "nj", "NI", "NJ", "NV", "ghost", "flux_method", "relax_type"):
        setattr(p, k, int(consts[k]))
    for k in ("Pr", "Kn_eff", "gam", "B", "K", "omega", "CFL", "umax", "vmax"):
        setattr(p, k, float(consts[k]))
    for k in ("wall_D", "wall_U", "wall_V", "wall_T", "wall_P", "beta_n", "beta_t", "gamma_f", "alpha_p", "adsorb",
              "alpha_n", "alpha_t"):
        setattr(p, k, (D * 4)(*[float(x) for x in consts[k]]))
    for k in ("dynamic_vartheta", "constant_adsorb"):
        setattr(p, k, (I * 4)(*[int(x) for x in consts[k]]))
    p.quad, p.weight = q.ctypes.data, w.ctypes.data
    p.mirror_NS, p.mirror_EW, p.mirror_D = mns.ctypes.data, mew.ctypes.data, md.ctypes.data
    for k, (pts, tri, nbr) in enumerate(zip(consts["iso"], consts["iso_tri"], consts["iso_nbr"])):
        if pts is not None:      # the arrays stay alive in consts
            p.iso[k], p.iso_tri[k], p.iso_nbr[k] = pts.ctypes.data, tri.ctypes.data, nbr.ctypes.data
            p.n_iso[k], p.n_tri[k] = len(pts), len(tri)
    return p


class OracleBlock:
    def __init__(self, gdata, bdef, xy):
        from pyugks_amd.data import initial_macro
        self.consts = gdata.block_constants(bdef)
        self.tables = gdata.velocity_tables()
        c = self.consts
        NI, NJ, NV, ni, nj = c["NI"], c["NJ"], c["NV"], c["ni"], c["nj"]
        wl = max(ni, nj)
        z = np.zeros
        self.arrays = dict(
            f=z((NI, NJ, NV, 2)), flux_f_S=z((NI, NJ, NV, 2)), flux_f_W=z((NI, NJ, NV, 2)),
            sigma=z((NI, NJ, NV, 2)), macro=z((ni, nj, 4)), Q=z((ni, nj, 2)), flux_macro_S=z((NI, NJ, 4)),
            flux_macro_W=z((NI, NJ, 4)), prim=z((NI, NJ, 4)), aL=z((NI, NJ, 4)), aR=z((NI, NJ, 4)),
            aT=z((NI, NJ, 4)), faceQ=z((NI, NJ, 2)), Mxi=z((NI, NJ, 3)),
            xy=np.ascontiguousarray(xy, dtype=np.float64), area=np.full((NI, NJ), -1.0),
            centre=np.full((NI, NJ, 2), -1.0), mid_side=np.full((NI, NJ, 2, 2), -1.0),
            normal=np.full((NI, NJ, 2, 2), -1.0), side_length=np.full((NI, NJ, 2), -1.0),
            wall_prop=z((wl, 4, 4)), wall_cover=z((wl, 4, 4)), wall_fluxes=z((wl, 4, 8)),
            wall_dist=z((wl, NV, 2)), time_step=z((ni, nj)), residual=z((ni, nj, 4)), bc_mdot=z((wl, 4)),
            flag=z(2))
        m, q = initial_macro(bdef)
        self.arrays["macro"][:] = m
        self.arrays["Q"][:] = q
        for bci, bcx in enumerate(bdef.bc_list):
            self.arrays["wall_cover"][:, bci, 0] = bcx.cover_initial
        self.block = Block()
        self.block.p = make_params(c, self.tables)
        st = State()
        for n in STATE_FIELDS:
            setattr(st, n, self.arrays[n].ctypes.data)
        self.block.s = st
        rt = {"S": 0, "D": 1, "CL": 2}
        for k, bcx in enumerate(bdef.bc_list):
            self.block.bc[k] = BC(int(bcx.type_of_BC), int(bcx.other_block), int(bcx.other_face),
                                  int(bcx.orientation), int(bcx.flip_distribution), rt.get(bcx.reflect_type, 0))

    def __getitem__(self, name):
        return self.arrays[name]

    def internal(self):
        """getInternal (clKernel.c:315-372): (Tx, Ty, Tz, T) and (tau_xx,
        tau_xy, tau_yy, 0) per interior cell of the block's current f."""
        c = self.consts
        txyz, stress = np.zeros((c["ni"], c["nj"], 4)), np.zeros((c["ni"], c["nj"], 4))
        lib().or_getInternal(C.byref(self.block.p), self.arrays["f"].ctypes.data, txyz.ctypes.data,
                             stress.ctypes.data)
        return txyz, stress


# ---- wall values given as expressions (UDF_D/U/V/T), restated ------------
# The reference pastes each UDF string into the OpenCL source as a #define
# (ugks_CL.py:278-298) and paraBC evaluates it per wall cell in s, x, y, t
# (clKernel.c:141-218).  Here the string is walked as a C expression tree with
# numpy: integer literals stay integers (C division and remainder truncate
# toward zero), anything touching a double is double; PI is '%0.15e' % pi.
_UDF_FUNCS = {"sin": np.sin, "cos": np.cos, "tan": np.tan, "exp": np.exp, "log": np.log, "sqrt": np.sqrt,
              "fabs": np.abs, "tanh": np.tanh, "atan": np.arctan, "asin": np.arcsin, "acos": np.arccos,
              "sinh": np.sinh, "cosh": np.cosh, "log10": np.log10, "floor": np.floor, "ceil": np.ceil,
              "pow": np.power, "fmin": np.fmin, "fmax": np.fmax, "atan2": np.arctan2, "fmod": np.fmod}


def udf_values(text: str, s, x, y, t):
    """The value of a UDF string at wall cells (s, x, y arrays) and time t."""
    import ast

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant):
            return n.value, isinstance(n.value, int)
        if isinstance(n, ast.Name):
            env = {"s": s, "x": x, "y": y, "t": t, "PI": 3.141592653589793, "M_PI": 3.141592653589793}
            return np.asarray(env[n.id], dtype=np.float64), False
        if isinstance(n, ast.UnaryOp):
            v, k = ev(n.operand)
            return (-v if isinstance(n.op, ast.USub) else v), k
        if isinstance(n, ast.BinOp):
            (a, ka), (b, kb) = ev(n.left), ev(n.right)
            if ka and kb:
                if isinstance(n.op, (ast.Div, ast.Mod)):
                    q = abs(a) // abs(b) * (1 if (a >= 0) == (b >= 0) else -1)
                    return (q if isinstance(n.op, ast.Div) else a - q * b), True
                return {ast.Add: a + b, ast.Sub: a - b, ast.Mult: a * b}[type(n.op)], True
            a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
            with np.errstate(all="ignore"):
                return {ast.Add: lambda: a + b, ast.Sub: lambda: a - b, ast.Mult: lambda: a * b,
                        ast.Div: lambda: a / b}[type(n.op)](), False
        if isinstance(n, ast.Call):
            args = [np.asarray(ev(a)[0], dtype=np.float64) for a in n.args]
            with np.errstate(all="ignore"):
                return _UDF_FUNCS[n.func.id](*args), False
        raise ValueError(f"unsupported UDF syntax: {ast.dump(n)[:60]}")

    v = np.asarray(ev(ast.parse(text.strip(), mode="eval"))[0], dtype=np.float64)
    return np.broadcast_to(v, np.shape(s)).astype(np.float64)


def wall_para(ni, nj, G, side_length, x, y):
    """(s, x, y) of every wall cell, [max(ni, nj)][4][3] (ugks_block.py:776-836):
    s the cell midpoint's arc length over the wall length, (x, y) the mean of
    its two vertices, the west wall's taken from vertex column 1 (:786)."""
    out = np.zeros((max(ni, nj), 4, 3))
    L = side_length
    for w, lens, xs, ys in ((0, L[G:G + ni, G + nj, 0], x[:, nj], y[:, nj]), (1, L[G + ni, G:G + nj, 1], x[ni, :], y[ni, :]),
                            (2, L[G:G + ni, G, 0], x[:, 0], y[:, 0]), (3, L[G, G:G + nj, 1], x[1, :], y[1, :])):
        cum = np.concatenate([[0.0], np.cumsum(lens)])
        out[:len(lens), w, 0] = (cum[:-1] + 0.5 * lens) / cum[-1]
        out[:len(lens), w, 1] = 0.5 * (xs[:-1] + xs[1:])
        out[:len(lens), w, 2] = 0.5 * (ys[:-1] + ys[1:])
    return out


class OracleSim:
    """The reference sequence (UGKSim.__init__ step part, one_step, run) in C."""

    def __init__(self, gdata, prepared: bool = False, threads: int | None = None):
        from pyugks_amd.data import ghost_coordinates
        if not prepared:
            gdata.prepare()
        self.gdata = gdata
        if threads is not None:
            set_threads(threads)
        xys = ghost_coordinates(gdata.blocks, gdata.ghost)
        self.blocks = [OracleBlock(gdata, b, xys[i]) for i, b in enumerate(gdata.blocks)]
        n = len(self.blocks)
        self._barr = (Block * n)(*[b.block for b in self.blocks])
        self.max_dt = np.zeros(n)
        self.sim = Sim()
        self.sim.nblocks = n
        self.sim.blocks = C.cast(self._barr, C.POINTER(Block))
        self.sim.dt = gdata.dt
        self.sim.time = gdata.time
        self.sim.CFL = gdata.CFL
        self.sim.delta_dt = gdata.delta_dt
        self.sim.suggest_dt = gdata.suggest_dt if gdata.suggest_dt else 0.0
        self.sim.Kn_eff = gdata.Kn_eff
        self.sim.step = 0
        self.sim.force_time = 0
        self.sim.dt_update_count = gdata.dt_update_count
        self.sim.max_dt = self.max_dt.ctypes.data
        # UGKSim.__init__ from initFunctions on (ugks_sim.py:81-92), with the
        # wall-value expressions evaluated where paraBC runs
        L = lib()
        self.udf = [[bc if getattr(bc, "UDF_D", None) is not None or getattr(bc, "UDF_U", None) is not None
                     or getattr(bc, "UDF_V", None) is not None or getattr(bc, "UDF_T", None) is not None else None
                     for bc in bd.bc_list] for bd in gdata.blocks]
        self.time_dependent = [any(u is not None and "t" in u for bc in bd.bc_list
                                   for u in (getattr(bc, "UDF_D", None), getattr(bc, "UDF_T", None),
                                             getattr(bc, "UDF_U", None), getattr(bc, "UDF_V", None)))
                               for bd in gdata.blocks]
        for b in self.blocks:
            a = b.arrays
            L.or_cellGeom(C.byref(b.block.p), *[C.c_void_p(a[k].ctypes.data) for k in
                                                 ("xy", "area", "centre", "mid_side", "normal", "side_length")])
        for bi, b in enumerate(self.blocks):
            L.or_block_init(C.byref(self._barr[bi]))
            self._para_udf(bi)
        L.or_sim_updateAllBC(C.byref(self.sim))
        L.or_sim_calcTimeStep(C.byref(self.sim))
        L.or_sim_updateTimeStep(C.byref(self.sim), 0)

    def _para_udf(self, bi):
        """paraBC on the walls with UDF strings (clKernel.c:141-218): D, U, V
        and T from the expressions (or the constants), s3 = P / D with a
        pressure, else 1 / T."""
        bcs = self.udf[bi]
        if not any(x is not None for x in bcs):
            return
        b, bd = self.blocks[bi], self.gdata.blocks[bi]
        c = b.consts
        ni, nj, G = c["ni"], c["nj"], c["ghost"]
        para = wall_para(ni, nj, G, b["side_length"], np.asarray(bd.x, dtype=np.float64),
                         np.asarray(bd.y, dtype=np.float64))
        t = self.sim.time
        for w, bc in enumerate(bcs):
            if bc is None or bc.type_of_BC not in (2, 4, 6, 7, 8):     # CONSTANT DIFFUSE INFLOW OUTFLOW ADSORBING
                continue
            n = ni if w in (0, 2) else nj
            s_, x_, y_ = para[:n, w, 0], para[:n, w, 1], para[:n, w, 2]

            def val(name, k):
                text = getattr(bc, name, None)
                return udf_values(text, s_, x_, y_, t) if text is not None else np.full(n, c[k][w])
            D_ = val("UDF_D", "wall_D")
            wp = b["wall_prop"]
            wp[:n, w, 0] = D_
            wp[:n, w, 1] = val("UDF_U", "wall_U")
            wp[:n, w, 2] = val("UDF_V", "wall_V")
            if c["wall_P"][w] != -1:
                wp[:n, w, 3] = c["wall_P"][w] / D_
            elif bc.UDF_T is not None:
                wp[:n, w, 3] = 1.0 / udf_values(bc.UDF_T, s_, x_, y_, t)
            else:
                wp[:n, w, 3] = c["wall_T"][w]

    def one_step(self, get_dt=True):
        for bi, td in enumerate(self.time_dependent):
            if td:
                self._para_udf(bi)          # UGKSBlock.parametricBC at the step's start time
        rc = lib().or_sim_one_step(C.byref(self.sim), int(bool(get_dt)))
        if rc:
            raise RuntimeError(f"oracle step failed {rc}")

    def run(self, nsteps: int):
        if any(self.time_dependent):
            for _ in range(int(nsteps)):
                self.one_step(self._get_dt())
            return
        rc = lib().or_sim_run(C.byref(self.sim), int(nsteps))
        if rc:
            raise RuntimeError(f"oracle run failed {rc}")

    def _get_dt(self):
        """or_sim_run's dt cadence (ugks_sim.py:542-578)."""
        n, st = self.sim.dt_update_count, self.sim.step
        return (st == -n) if n < 0 else ((st + 1) % n == 0)

    @property
    def dt(self):
        return self.sim.dt

    @property
    def time(self):
        return self.sim.time

    def macro(self):
        return [b["macro"].copy() for b in self.blocks]
