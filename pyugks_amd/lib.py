"""ctypes binding of libugks_hip.so (declared in include/ugks_hip.h).

The library is the only compute path: if it is missing or cannot be loaded,
:func:`load` raises; nothing falls back to host code.

``Program`` plays the role of the pyopencl program object the reference
keeps per block (``self.prg``, ugks_block.py:184): ``prg.iFace(stream,
Fin, iface_f, ...)`` launches the HIP kernel that replaces the OpenCL kernel
of the same name, with the reference's argument order.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libugks_hip.so")

P = C.c_void_p
I = C.c_int32
D = C.c_double

# name -> argtypes after (blk, stream); reference kernel names (source/*.c)
KERNELS = {
    "cellGeom": [P, P, P, P, P, P],
    "paraBC": [P, P, D],
    "clFindDT": [P, P, P, P, D],
    "time_step_max": [P, P],
    "initFunctions": [P, P, P],
    "calcMacro": [P, P],
    "calcQ": [P, P, P],
    "getInternal": [P, P, P],
    "iFace": [P, P, P, P, P, P, I],
    "getFaceCons": [P, P, I, P, I, I],
    "getAL": [P, P, I, P, I, I],
    "getAR": [P, P, I, P, I, I],
    "getPLR": [P, P, I, P, P, P, I, I],
    "initMacroFlux": [P, I, D, P, P, P, P, P, I, I],
    "calcFaceQ": [P, P, P, P, I, I, I],
    "macroFlux": [P, P, P, P, P, I, D, P, P, I, I],
    "distFlux": [P, P, P, P, I, D, P, P, P, P, P, P, I, I],
    "accommodatingWallDist": [P, I, P, P, D],
    "wallFlux": [P, P, I, P, P, D],
    "wallMassEnergyFluxes": [P, P, P, I, P, P, D],
    "adsorbingWall_P1": [P, I, P, P, P, P, P, D, P],
    "adsorbingWallS_P2": [P, I, P, P, P, P, P, D],
    "adsorbingWallD_P2": [P, I, P, P, P, P, P, D],
    "adsorbingWallCL_P2": [P, I, P, P, P, P, P, D],
    "updateMacro": [P, P, P, P, P],
    "UGKS_update": [P, P, P, P, P, P, P, D, P],
    "getResidual": [P, P],
    "edgeMassFlow": [P, P, P, P, I],
    "edgeExchange": [P, P, I, P, P, I, I, I, I, I],
    "edgeExtrapolate": [P, I],
    "edgeConstant": [P, I, P],
    "edgeInflow": [P, I, P, P, P, P],
    "edgeOutflow": [P, I, P, P, P, P],
    "edgeMirror": [P, I],
    "edgeBounceBack": [P, I],
    "edgeConstGrad": [P, I],
}

# fused / batched entry points: name -> argtypes (full list)
FUSED = {
    "ugks_step_plan_create": [P, P, P, I, I, P, C.POINTER(P)],
    "ugks_step_plan_create_batch": [P, P, P, I, P, I, I, P, C.POINTER(P)],
    "ugks_step_plan_destroy": [P],
    "ugks_step_run": [P, P, I],
    "ugks_step_run_graph": [P, P, I],
    "ugks_step_autotune": [P, P, I, P],
    "ugks_step_run_timed": [P, P, I, P],
    "ugks_step_state": [P, P, P, P],
    "ugks_step_states": [P, P, P, P],
    "ugks_step_set_dt": [P, P, D],
    "ugks_step_set_dt_case": [P, P, I, D],
    "ugks_step_force_dt_case": [P, P, I, D],
    "ugks_step_block_fits": [I, I, I, I],
    "ugks_copy_kernel": [P, P, P, C.c_int64],
    "ugks_f64_to_f32": [P, P, P, C.c_int64],
    "ugks_f32_to_f64": [P, P, P, C.c_int64],
    "ugks_step_run_noalloc": [P, P, P, I, I, P, I, P],
}
UGKS_FP64, UGKS_MIXED = 0, 1


class BlockBuffers(C.Structure):
    """ugks_block_buffers (include/ugks_hip.h)"""
    _fields_ = [(n, P) for n in (
        "f", "flux_f_S", "flux_f_W", "flux_macro_S", "flux_macro_W", "macro", "macro_new", "residual", "Q",
        "xy", "centre", "mid_side", "normal", "area", "side_length", "wall_prop", "wall_cover", "wall_fluxes",
        "wall_dist", "time_step", "flag", "para")]


class BCDesc(C.Structure):
    _fields_ = [("type", I), ("other_block", I), ("other_face", I), ("orientation", I), ("flip", I),
                ("reflect_type", I)]


class DtDesc(C.Structure):
    _fields_ = [("dt", D), ("time", D), ("CFL", D), ("delta_dt", D), ("suggest_dt", D), ("Kn_eff", D),
                ("step", I), ("overshoot", I), ("force_time", I), ("pad", I), ("max_dt", D)]

RUNTIME = {
    "ugks_block_create": [P, C.POINTER(P)],
    "ugks_block_destroy": [P],
    "ugks_get_device_count": [C.POINTER(C.c_int)],
    "ugks_set_device": [C.c_int],
    "ugks_malloc": [C.POINTER(P), C.c_size_t],
    "ugks_free": [P],
    "ugks_memset": [P, C.c_int, C.c_size_t, P],
    "ugks_memcpy_htod": [P, P, C.c_size_t, P],
    "ugks_memcpy_dtoh": [P, P, C.c_size_t, P],
    "ugks_memcpy_dtod": [P, P, C.c_size_t, P],
    "ugks_stream_create": [C.POINTER(P)],
    "ugks_stream_destroy": [P],
    "ugks_stream_synchronize": [P],
    "ugks_event_create": [C.POINTER(P)],
    "ugks_event_destroy": [P],
    "ugks_event_record": [P, P],
    "ugks_event_elapsed_ms": [C.POINTER(C.c_float), P, P],
}


class BlockDesc(C.Structure):
    """ugks_block_desc (include/ugks_hip.h)"""
    _fields_ = [
        ("ni", I), ("nj", I), ("NI", I), ("NJ", I), ("NV", I), ("ghost", I),
        ("flux_method", I), ("relax_type", I),
        ("Pr", D), ("Kn_eff", D), ("gam", D), ("B", D), ("K", D), ("omega", D),
        ("CFL", D), ("umax", D), ("vmax", D),
        ("wall_D", D * 4), ("wall_U", D * 4), ("wall_V", D * 4), ("wall_T", D * 4), ("wall_P", D * 4),
        ("beta_n", D * 4), ("beta_t", D * 4), ("gamma_f", D * 4), ("alpha_p", D * 4),
        ("dynamic_vartheta", I * 4), ("constant_adsorb", I * 4),
        ("adsorb", D * 4),
        ("quad", C.POINTER(D)), ("weight", C.POINTER(D)),
        ("mirror_NS", C.POINTER(C.c_int64)), ("mirror_EW", C.POINTER(C.c_int64)),
        ("mirror_D", C.POINTER(C.c_int64)),
        ("alpha_n", D * 4), ("alpha_t", D * 4),
        ("iso", C.c_void_p * 4), ("iso_tri", C.c_void_p * 4), ("iso_nbr", C.c_void_p * 4),
        ("n_iso", I * 4), ("n_tri", I * 4),
        ("udf_code", C.c_void_p), ("udf_const", C.c_void_p), ("udf_ncode", I), ("udf_nconst", I),
        ("udf_start", I * 16), ("udf_len", I * 16),
    ]


class UGKSError(RuntimeError):
    pass


_lib = None


def load(path: str | None = None):
    """Load libugks_hip.so once; raise if it is absent (no host fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or os.environ.get("UGKS_HIP_LIB", LIB_PATH)
    if not os.path.exists(p):
        raise UGKSError(f"{p} is missing: build it with `python -m pyugks_amd.build` "
                        "(the UGKS step has no host fallback)")
    lib = C.CDLL(p)
    for name, argt in RUNTIME.items():
        f = getattr(lib, name)
        f.argtypes = argt
        f.restype = C.c_int
    for name, argt in KERNELS.items():
        f = getattr(lib, "ugks_" + name)
        f.argtypes = [P, P] + argt
        f.restype = C.c_int
    for name, argt in FUSED.items():
        f = getattr(lib, name)
        f.argtypes = argt
        f.restype = C.c_int
    lib.ugks_last_error.restype = C.c_char_p
    lib.ugks_last_error.argtypes = []
    lib.ugks_step_noalloc_capacity.restype = C.c_size_t
    lib.ugks_step_noalloc_capacity.argtypes = []
    _lib = lib
    return lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = load().ugks_last_error().decode(errors="replace")
        raise UGKSError(f"{what or 'libugks_hip'} failed ({rc}): {msg}")
    return rc


def exported_symbols():
    """Every C symbol include/ugks_hip.h declares."""
    return (list(RUNTIME) + ["ugks_last_error", "ugks_step_noalloc_capacity"] + ["ugks_" + k for k in KERNELS]
            + list(FUSED))


# ------------------------------------------------------------------ memory --
class DeviceArray:
    """A device allocation with a numpy shape/dtype view for transfers."""

    def __init__(self, shape, dtype=np.float64, fill=None):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        lib = load()
        ptr = P()
        check(lib.ugks_malloc(C.byref(ptr), self.nbytes), "ugks_malloc")
        self.ptr = ptr.value
        if fill is not None:
            if fill == 0:
                self.zero()
            else:
                self.upload(np.full(self.shape, fill, dtype=self.dtype))

    @classmethod
    def from_host(cls, arr, dtype=None):
        a = np.ascontiguousarray(arr, dtype=dtype or np.asarray(arr).dtype)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, arr, stream=None):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"size mismatch {a.nbytes} != {self.nbytes}")
        check(load().ugks_memcpy_htod(self.ptr, a.ctypes.data, self.nbytes, stream), "upload")

    def download(self, stream=None):
        out = np.empty(self.shape, dtype=self.dtype)
        check(load().ugks_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes, stream), "download")
        return out

    def copy_from(self, other: "DeviceArray", stream=None):
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        check(load().ugks_memcpy_dtod(self.ptr, other.ptr, self.nbytes, stream), "dtod")

    def zero(self, stream=None):
        check(load().ugks_memset(self.ptr, 0, self.nbytes, stream), "memset")

    def free(self):
        if getattr(self, "ptr", None):
            load().ugks_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def _as_parameter_(self):
        return P(self.ptr)


def _arg(x):
    if isinstance(x, DeviceArray):
        return x.ptr
    return x


class Program:
    """Per-block kernel set: ``prg.<reference kernel name>(stream, *args)``."""

    def __init__(self, desc: BlockDesc, keep=()):
        lib = load()
        h = P()
        check(lib.ugks_block_create(C.byref(desc), C.byref(h)), "ugks_block_create")
        self.handle = h.value
        self._keep = keep   # host tables referenced by desc stay alive until create returns

    def __getattr__(self, name):
        if name not in KERNELS:
            raise AttributeError(name)
        f = getattr(load(), "ugks_" + name)

        def call(stream, *args):
            return check(f(self.handle, stream, *[_arg(a) for a in args]), name)

        call.__name__ = name
        return call

    def close(self):
        if getattr(self, "handle", None):
            load().ugks_block_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CLCompatProgram(Program):
    """The reference's call shape, unchanged: ``prg.<kernel>(queue, global_size,
    local_size, *args)`` (ugks_block.py:1066-1295).  ``queue`` is a HIP stream
    handle (or None); the NDRange sizes are accepted and ignored, because each
    kernel derives its launch from the block and its face/offset arguments;
    buffers are DeviceArray objects or raw device pointers."""

    def __getattr__(self, name):
        if name not in KERNELS:
            raise AttributeError(name)
        f = getattr(load(), "ugks_" + name)

        def call(queue, global_size, local_size, *args):
            return check(f(self.handle, queue, *[_arg(a) for a in args]), name)

        call.__name__ = name
        return call


def make_desc(consts: dict, tables) -> tuple:
    """BlockDesc from UGKSData.block_constants() and velocity_tables()."""
    q, w, mns, mew, md = tables
    d = BlockDesc()
    for k in ("ni", "nj", "NI", "NJ", "NV", "ghost", "flux_method", "relax_type"):
        setattr(d, k, int(consts[k]))
    for k in ("Pr", "Kn_eff", "gam", "B", "K", "omega", "CFL", "umax", "vmax"):
        setattr(d, k, float(consts[k]))
    for k in ("wall_D", "wall_U", "wall_V", "wall_T", "wall_P", "beta_n", "beta_t", "gamma_f", "alpha_p", "adsorb",
              "alpha_n", "alpha_t"):
        setattr(d, k, (D * 4)(*[float(x) for x in consts[k]]))
    for k in ("dynamic_vartheta", "constant_adsorb"):
        setattr(d, k, (I * 4)(*[int(x) for x in consts[k]]))
    d.quad = q.ctypes.data_as(C.POINTER(D))
    d.weight = w.ctypes.data_as(C.POINTER(D))
    d.mirror_NS = mns.ctypes.data_as(C.POINTER(C.c_int64))
    d.mirror_EW = mew.ctypes.data_as(C.POINTER(C.c_int64))
    d.mirror_D = md.ctypes.data_as(C.POINTER(C.c_int64))
    keep = [q, w, mns, mew, md]
    for k, (pts, tri, nbr) in enumerate(zip(consts["iso"], consts["iso_tri"], consts["iso_nbr"])):
        if pts is not None:
            d.iso[k], d.iso_tri[k], d.iso_nbr[k] = pts.ctypes.data, tri.ctypes.data, nbr.ctypes.data
            d.n_iso[k], d.n_tri[k] = len(pts), len(tri)
            keep += [pts, tri, nbr]
    udf = consts.get("udf")
    if udf is not None:
        code, cst, start, length = udf
        d.udf_code, d.udf_const = code.ctypes.data, cst.ctypes.data
        d.udf_ncode, d.udf_nconst = len(code), len(cst)
        d.udf_start = (I * 16)(*start)
        d.udf_len = (I * 16)(*length)
        keep += [code, cst]
    return d, tuple(keep)


class Event:
    def __init__(self):
        h = P()
        check(load().ugks_event_create(C.byref(h)), "event")
        self.h = h.value

    def record(self, stream=None):
        check(load().ugks_event_record(self.h, stream), "event record")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(load().ugks_event_elapsed_ms(C.byref(ms), self.h, end.h), "elapsed")
        return float(ms.value)

    def __del__(self):
        try:
            load().ugks_event_destroy(self.h)
        except Exception:
            pass


def synchronize(stream=None):
    check(load().ugks_stream_synchronize(stream), "synchronize")


def device_count() -> int:
    n = C.c_int(0)
    rc = load().ugks_get_device_count(C.byref(n))
    return n.value if rc == 0 else 0
