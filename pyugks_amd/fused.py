"""The fused, batched step (ugks_step_* in include/ugks_hip.h).

``StepPlan(sims, precision)`` takes one simulation, or a list of independent
simulations (cases), that the reference-shaped path has initialised
(geometry, initial f, ghosts, first time step; the sequence of
ugks_sim.py:52-92) and advances all of their blocks together: each step is
eight launches over every block of every case, with each case's time step
kept on the device (one UGKSim.one_step per case, ugks_sim.py:407-501).
With ``graph=True`` the step runs as replays of a captured HIP graph.  With
``precision="mixed"`` f and the per-velocity face fluxes are stored and
computed in FP32; moments, the macroscopic update, heat fluxes, equilibrium
moments and the time step stay FP64.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L


def plan_inputs(sims, precision, stream=None):
    """The C-ABI inputs of a step over the blocks of ``sims`` (cases): block
    handles, buffer records (with the FP32 copies of f and of the face fluxes
    allocated and filled for the mixed step), boundary records, the case of
    each block and one time-step record per case."""
    lib = L.load()
    mixed = precision == "mixed"
    blocks = [(ci, b) for ci, sim in enumerate(sims) for b in sim.blocks]
    first = np.cumsum([0] + [len(sim.blocks) for sim in sims])
    nb = len(blocks)
    bufs = (L.BlockBuffers * nb)()
    bcs = (L.BCDesc * (4 * nb))()
    handles = (C.c_void_p * nb)()
    case_of = (C.c_int32 * nb)(*[ci for ci, _ in blocks])
    rt = {"S": 0, "D": 1, "CL": 2}
    for bi, (ci, b) in enumerate(blocks):
        b.macro_new_D = L.DeviceArray((b.ni, b.nj, 4), fill=0.0)
        if mixed:
            dist = (b.Ni, b.Nj, b.Nv, 2)
            b.f32_D = L.DeviceArray(dist, np.float32, fill=0.0)
            b.fluxS32_D = L.DeviceArray(dist, np.float32, fill=0.0)
            b.fluxW32_D = L.DeviceArray(dist, np.float32, fill=0.0)
            n = int(np.prod(dist))
            L.check(lib.ugks_f64_to_f32(stream, b.f_D.ptr, b.f32_D.ptr, n), "f64_to_f32")
            f, fs, fw = b.f32_D, b.fluxS32_D, b.fluxW32_D
        else:
            f, fs, fw = b.f_D, b.flux_f_S_D, b.flux_f_W_D
        u = bufs[bi]
        u.f, u.flux_f_S, u.flux_f_W = f.ptr, fs.ptr, fw.ptr
        u.flux_macro_S, u.flux_macro_W = b.flux_macro_S_D.ptr, b.flux_macro_W_D.ptr
        u.macro, u.macro_new, u.residual, u.Q = b.macro_D.ptr, b.macro_new_D.ptr, b.residual_D.ptr, b.Q_D.ptr
        u.xy, u.centre, u.mid_side, u.normal = b.xy_D.ptr, b.centre_D.ptr, b.side_D.ptr, b.normal_D.ptr
        u.area, u.side_length = b.area_D.ptr, b.length_D.ptr
        u.wall_prop, u.wall_cover, u.wall_fluxes = b.wall_prop_D.ptr, b.wall_cover_D.ptr, b.wall_fluxes_D.ptr
        u.wall_dist, u.time_step, u.flag = b.wall_dist_D.ptr, b.time_step_D.ptr, b.flag_D.ptr
        u.para = b.para_D.ptr
        handles[bi] = b.prg.handle
        for k, x in enumerate(b.bc_list):
            other = int(x.other_block)
            if other >= 0:
                other += int(first[ci])
            bcs[4 * bi + k] = L.BCDesc(int(x.type_of_BC), other, int(x.other_face),
                                       int(x.orientation), int(x.flip_distribution), rt.get(x.reflect_type, 0))
    dts = (L.DtDesc * len(sims))()
    for ci, sim in enumerate(sims):
        g = sim.gdata
        # the blocks' maximum time steps of the last calcTimeStep: the first
        # step re-evaluates dt against them, as UGKSim.one_step does
        md = [float(getattr(b, "max_dt", 0.0) or 0.0) for b in sim.blocks]
        dts[ci] = L.DtDesc(g.dt, g.time, g.CFL, g.delta_dt, g.suggest_dt if g.suggest_dt else 0.0, g.Kn_eff,
                           int(sim.step), 0, int(bool(getattr(sim, "force_time", False))), 0,
                           min(md) if all(m > 0 for m in md) else 0.0)
    return blocks, bufs, bcs, handles, case_of, dts


class StepPlan:
    KERNELS = ("k_faceA", "k_faceS", "k_faceC", "k_wall", "k_update", "k_dt", "k_dtfin", "k_ghost")

    def __init__(self, sims, precision: str = "fp64", stream=None, graph: bool = False):
        self.single = not isinstance(sims, (list, tuple))
        self.sims = [sims] if self.single else list(sims)
        self.sim = self.sims[0]
        self.stream = stream
        self.precision = precision
        self.graph = graph
        mixed = precision == "mixed"
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        lib = L.load()
        blocks, bufs, bcs, handles, case_of, dts = plan_inputs(self.sims, precision, stream)
        nb = len(blocks)
        h = C.c_void_p()
        L.check(lib.ugks_step_plan_create_batch(handles, bufs, bcs, nb, case_of, len(self.sims),
                                                L.UGKS_MIXED if mixed else L.UGKS_FP64, dts, C.byref(h)),
                "ugks_step_plan_create_batch")
        self.handle = h.value
        self.blocks = [b for _, b in blocks]
        self._keep = (bufs, bcs, handles, case_of, dts)

    def run(self, nsteps: int):
        lib = L.load()
        if self.graph:
            L.check(lib.ugks_step_run_graph(self.handle, self.stream, int(nsteps)), "ugks_step_run_graph")
        else:
            L.check(lib.ugks_step_run(self.handle, self.stream, int(nsteps)), "ugks_step_run")

    def autotune(self, reps: int = 3) -> dict:
        """Pick the face passes' workgroup sizes by timing them (state unchanged)."""
        choice = (C.c_int32 * 2)()
        L.check(L.load().ugks_step_autotune(self.handle, self.stream, int(reps), choice), "ugks_step_autotune")
        return {"k_faceA_waves_per_workgroup": int(choice[0]), "k_faceC_waves_per_workgroup": int(choice[1])}

    def run_timed(self, nsteps: int) -> dict:
        """Run nsteps with HIP events around each launch; per-kernel total ms."""
        ms = (C.c_float * len(self.KERNELS))()
        L.check(L.load().ugks_step_run_timed(self.handle, self.stream, int(nsteps), ms), "ugks_step_run_timed")
        return dict(zip(self.KERNELS, [float(x) for x in ms]))

    def states(self, update_sim: bool = True):
        """Time-step state of every case (and the per-block maximum time steps)."""
        ds = (L.DtDesc * len(self.sims))()
        md = np.zeros(len(self.blocks))
        L.check(L.load().ugks_step_states(self.handle, self.stream, ds, md.ctypes.data_as(C.c_void_p)),
                "ugks_step_states")
        if update_sim:
            k = 0
            for sim, d in zip(self.sims, ds):
                g = sim.gdata
                g.dt, g.time, g.step = d.dt, d.time, d.step
                sim.step = d.step
                for b in sim.blocks:
                    b.max_dt = float(md[k])
                    k += 1
        return list(ds)

    def state(self, update_sim: bool = True):
        return self.states(update_sim)[0]

    def set_dt(self, dt: float, case: int = 0, hold: bool = True):
        """Set one case's time step; with ``hold`` it stays fixed for the
        remaining steps (the shortened last step onto max_time)."""
        fn = L.load().ugks_step_force_dt_case if hold else L.load().ugks_step_set_dt_case
        L.check(fn(self.handle, self.stream, int(case), float(dt)), "set dt")

    def f64(self, block):
        """f of one block as FP64 (converted from FP32 for the mixed step)."""
        if self.precision == "mixed":
            out = L.DeviceArray(block.f32_D.shape, np.float64)
            L.check(L.load().ugks_f32_to_f64(self.stream, block.f32_D.ptr, out.ptr, int(np.prod(out.shape))), "f32_to_f64")
            return out.download(self.stream)
        return block.f_D.download(self.stream)

    def flag(self):
        return max(float(b.flag_D.download(self.stream)[0]) for b in self.blocks)

    def close(self):
        if getattr(self, "handle", None):
            L.load().ugks_step_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_noalloc(sim, precision: str, nsteps: int, dt_dev=None, stream=None):
    """Advance one case by nsteps through ugks_step_run_noalloc, the step that
    allocates no device memory (its working memory is the library's static
    per-GPU array).  The case's buffers and, for the mixed step, the FP32
    copies of f are prepared beforehand (here, outside the call); ``dt_dev``
    is a device ugks_dt_desc, allocated here when not given.  Returns it; the
    simulation's dt, time and step are updated from it."""
    if precision not in ("fp64", "mixed"):
        raise ValueError("precision must be 'fp64' or 'mixed'")
    lib = L.load()
    blocks, bufs, bcs, handles, case_of, dts = plan_inputs([sim], precision, stream)
    if dt_dev is None:
        dt_dev = L.DeviceArray((C.sizeof(L.DtDesc),), np.uint8)
    dt_dev.upload(np.frombuffer(bytes(dts[0]), dtype=np.uint8), stream)
    L.check(lib.ugks_step_run_noalloc(handles, bufs, bcs, len(blocks), L.UGKS_MIXED if precision == "mixed" else
                                      L.UGKS_FP64, dt_dev.ptr, int(nsteps), stream), "ugks_step_run_noalloc")
    d = L.DtDesc.from_buffer_copy(dt_dev.download(stream).tobytes())
    g = sim.gdata
    g.dt, g.time, g.step = d.dt, d.time, d.step
    sim.step = d.step
    if precision == "mixed":
        for _, b in blocks:
            L.check(lib.ugks_f32_to_f64(stream, b.f32_D.ptr, b.f_D.ptr, int(np.prod(b.f32_D.shape))), "f32_to_f64")
    return d
