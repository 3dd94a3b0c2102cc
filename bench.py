#!/usr/bin/env python3
"""Benchmark of the UGKS time step on one MI355X.

Workload (default): the lid-driven cavity split into 2x2 blocks of 128x128
cells (256x256 in total) with ghost exchange between the blocks, 33x33
Newton-Cotes velocities on [-6,6]^2 (the "32x32" grid after the reference's
n -> floor(n/4)*4+1 rounding), WENO5, Kn = 0.01, lid 0.15, CFL 0.8, FP64.
Synthetic input: the initial state is the Maxwellian at rho = 1, U = 0,
T = 1 (no data files).

A step is one full time step of all four blocks (flux, update, time step,
ghost exchange) through the fused FP64 path; the metric is
cell-velocity updates per second = sum(ni*nj*Nv) * steps / seconds, with
all state resident in HBM.  The same workload is also run with the mixed
FP32/FP64 step and its error against the FP64 step is reported, and the
FP64 CPU oracle is timed on a bounded sample on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import copy
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell·velocity updates/s + % HBM roofline, mixed vs FP64 at ≤1e-5 rel err"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=128, help="cells per block side (2x2 blocks)")
    ap.add_argument("--nc", type=int, default=32, help="Newton-Cotes points per axis before rounding")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--profile-steps", type=int, default=5, help="steps timed per kernel with HIP events")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel on MIB MiB (byte-counter calibration) and exit")
    return ap.parse_args()


def calibrate_copy(mib, reps=5):
    from pyugks_amd import lib as L
    import numpy as np
    n = mib * (1 << 20)
    src = L.DeviceArray((n // 8,), np.float64, fill=0.0)
    dst = L.DeviceArray((n // 8,), np.float64, fill=0.0)
    lib = L.load()
    L.check(lib.ugks_copy_kernel(None, src.ptr, dst.ptr, n), "copy")
    L.synchronize()
    e0, e1 = L.Event(), L.Event()
    e0.record()
    for _ in range(reps):
        L.check(lib.ugks_copy_kernel(None, src.ptr, dst.ptr, n), "copy")
    e1.record()
    ms = e0.elapsed_ms(e1) / reps
    print(json.dumps({"copy_bytes": n, "read_plus_write_bytes": 2 * n, "ms": ms,
                      "gbs": 2 * n / (ms * 1e-3) / 1e9}), flush=True)


def host_cores():
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, min(n, 16))


def read_traffic(kernel, workload):
    """HBM bytes per launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE),
    from the summary committed under profiles/ for this kernel and workload."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) >= 3 and parts[0] == kernel and parts[1] == workload:
                try:
                    return float(parts[2])
                except ValueError:
                    return None
    return None


def timed_run(plan, steps, warmup):
    from pyugks_amd import lib as L
    plan.run(warmup)
    L.synchronize(plan.stream)
    e0, e1 = L.Event(), L.Event()
    t0 = time.perf_counter()
    e0.record(plan.stream)
    plan.run(steps)
    e1.record(plan.stream)
    L.synchronize(plan.stream)
    t1 = time.perf_counter()
    return t1 - t0, e0.elapsed_ms(e1) / 1e3


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: runs on one GPU only; --gpus {a.gpus} is not supported (blocks would be "
              "sharded one group per GPU with RCCL ghost exchange, which is not built)", file=sys.stderr)
        sys.exit(2)
    from pyugks_amd import cases
    from pyugks_amd import lib as L
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    L.load()
    if L.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return

    g0 = cases.cfg2_cavity_2x2(n_per_block=a.cells, nc=a.nc)
    workload = f"cavity_2x2x{a.cells}sq_nc{a.nc}_weno5"
    print(f"[bench] init {workload}", file=sys.stderr, flush=True)
    sim = UGKSim(copy.deepcopy(g0))
    nv = sim.gdata.Nv
    upd_per_step = sum(b.ni * b.nj for b in sim.blocks) * nv
    plan = StepPlan(sim, "fp64")
    wall_s, ev_s = timed_run(plan, a.steps, a.warmup)
    value = upd_per_step * a.steps / wall_s
    print(f"[bench] fp64 {a.steps} steps in {wall_s:.4f} s ({ev_s:.4f} s by events)", file=sys.stderr, flush=True)
    kern = plan.run_timed(a.profile_steps)
    per_launch = {k: v / a.profile_steps for k, v in kern.items()}
    dom = max(per_launch, key=per_launch.get)
    b_alg = 32.0  # (1 read + 1 write) x (g, h) x 8 B per cell-velocity update
    achieved = b_alg * upd_per_step / (per_launch[dom] * 1e-3) / 1e9
    traffic = read_traffic(dom, workload)
    macro64 = [b.macro() for b in sim.blocks]
    steps64 = plan.state().step
    flag64 = plan.flag()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-velocity updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * wall_s / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Maxwellian initial state, no input files)",
        "config": {"workload": workload, "blocks": len(sim.blocks), "cells": upd_per_step // nv, "Nv": nv,
                   "reconstruction": "WENO5", "Kn": 0.01, "lid": 0.15, "CFL": 0.8, "path": "fused fp64"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                     "kernel_ms": per_launch[dom], "b_alg_per_update": b_alg,
                     "step_achieved_gbs": value * b_alg / 1e9},
        "kernel_ms_per_launch": per_launch,
        "flag": flag64,
    }

    if not a.no_mixed:
        from pyugks_amd.metrics import field_errors
        simm = UGKSim(copy.deepcopy(g0))
        planm = StepPlan(simm, "mixed")
        wall_m, _ = timed_run(planm, a.steps, a.warmup)
        vm = upd_per_step * a.steps / wall_m
        km = planm.run_timed(a.profile_steps)
        plm = {k: v / a.profile_steps for k, v in km.items()}
        domm = max(plm, key=plm.get)
        # compare after the same number of steps as the FP64 state above
        extra = steps64 - planm.state().step
        if extra > 0:
            planm.run(extra)
        errs = field_errors([b.macro() for b in simm.blocks], macro64)
        ach_m = 16.0 * upd_per_step / (plm[domm] * 1e-3) / 1e9
        out["mixed"] = {"value": vm, "ms_per_step": 1e3 * wall_m / a.steps, "dtype": "f32 storage/flux, f64 moments",
                        "speedup_vs_fp64": vm / value,
                        "roofline": {"bound": "hbm", "achieved": ach_m, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                     "frac": ach_m / HBM_PEAK_GBS, "kernel": domm, "kernel_ms": plm[domm],
                                     "b_alg_per_update": 16.0, "step_achieved_gbs": vm * 16.0 / 1e9},
                        "err_vs_fp64": errs, "kernel_ms_per_launch": plm, "flag": planm.flag()}
        del planm, simm

    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        cores = host_cores()
        print(f"[bench] cpu baseline on {cores} threads", file=sys.stderr, flush=True)
        ora = O.OracleSim(copy.deepcopy(g0), threads=cores)
        t0 = time.perf_counter()
        ora.run(1)
        one = time.perf_counter() - t0
        n = max(1, int(a.cpu_seconds / max(one, 1e-3)) - 1)
        t0 = time.perf_counter()
        ora.run(n)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": upd_per_step * n / dt, "unit": "cell-velocity updates/s", "cores": cores,
                               "kind": "port",
                               "sample": f"{n} steps of the same workload (after 1 calibration step), FP64 "
                                         f"C/OpenMP oracle, {dt:.1f} s"}
        out["cpu_baseline"]["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
