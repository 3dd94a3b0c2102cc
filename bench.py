#!/usr/bin/env python3
"""Benchmark of the UGKS time step on one MI355X.

Default workload (batch64): 64 independent lid-driven cavities of uneven
size in ONE plan, drawn from a fixed seed: cells per side from {32, 64, 96,
128, 192, 256}, Newton-Cotes velocity grids "16/24/32/40" per axis (17^2,
25^2, 33^2, 41^2 after the reference's n -> floor(n/4)*4+1 rounding) on
[-6,6]^2, Kn log-uniform in [1e-3, 1], lid speed in [0.05, 0.3], wall T in
[0.8, 1.2], WENO5, CFL 0.8.  It is the mixed FP32/FP64 batched step the
metric is quoted on; the same batch also runs in plain FP64 beside it.
Synthetic input: each case starts from the Maxwellian at rho = 1, U = 0,
T = 1 (no data files).  Other workloads (--workload): cavity_2x2 (2x2 blocks
of 128^2 cells, 33^2 velocities, WENO5, Kn 0.01), cavity_64 (64^2,
Gauss-Hermite 24^2, vanLeer, Kn 0.075), cavity_512 (512^2, Newton-Cotes
49^2, rho perturbed by 1%, --kn), channel (adsorbing-wall channel, 8
blocks, 1024x256, 41^2 velocities).

A step is one full time step of every block of every case (flux, update,
time step, ghost exchange) through the fused path, replayed as a HIP graph
(--no-graph: eager launches); the metric is cell-velocity updates per
second = sum(ni*nj*Nv) * steps / seconds, with all state resident in HBM.
`value` is the mixed-precision rate; the FP64 rate, the mixed step's error
against the FP64 step (L-inf and L2 of rho, U, T, q) and the mass drift of
both are reported beside it.  The FP64 CPU oracle is timed on a bounded
sample of the same workload (all host threads of this process's CPU share,
and one thread; median of three runs each), with the FP64 GPU state
compared to it after the same steps.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import copy
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell·velocity updates/s + % HBM roofline, mixed vs FP64 at ≤1e-5 rel err"
HBM_PEAK_GBS = 8000.0
CUS, CLOCK_HZ = 256, 2.4e9      # MI355X: one full-rate wave64 VALU instruction per CU per clock
WORKLOADS = ("cavity_2x2", "cavity_64", "cavity_512", "batch64", "channel")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 5 warm-up + 20 timed + 5 per-kernel steps stay inside the batch's stable
    # horizon: its most fragile case (256^2 cells, 17^2 velocities, Kn 0.029)
    # meets the WENO5 scheme's instability near step 70 at CFL 0.8, and the
    # mixed step's U error there passes 1e-5 after about 45 steps
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=WORKLOADS, default="batch64")
    ap.add_argument("--cells", type=int, default=128, help="cavity_2x2: cells per block side")
    ap.add_argument("--nc", type=int, default=32, help="cavity_2x2: Newton-Cotes points per axis before rounding")
    ap.add_argument("--kn", type=float, default=0.075, help="cavity_512: Knudsen number")
    ap.add_argument("--precision", choices=("both", "fp64", "mixed"), default="both")
    ap.add_argument("--no-graph", action="store_true", help="time eager launches instead of graph replays")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=5.0, help="target length of one CPU sample run")
    ap.add_argument("--profile-steps", type=int, default=5, help="steps timed per kernel with HIP events")
    ap.add_argument("--per-kernel-steps", type=int, default=0, metavar="N",
                    help="also time N FP64 steps of the reference-shaped per-kernel path (UGKSim.one_step)")
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
    """Threads for the all-core CPU sample: the CPUs this process may run on,
    held to OMP_NUM_THREADS when that is set (the host's CPU share for one
    GPU); also the machine's nproc."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n), os.cpu_count() or 1


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def read_profile_table(fname, workload):
    """Per-launch figures {kernel: value} for one workload from a counter
    summary committed under profiles/ (lines "kernel workload value ..."), and
    the source commit the summary says it was measured on ("# source commit
    <hash>").  These are NOT measured by this run: they are labelled with
    their file and commit wherever they are reported."""
    path = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(path):
        return {}, None
    out, commit = {}, None
    with open(path) as fh:
        for line in fh:
            if line.startswith("# source commit"):
                commit = line.split()[-1]
                continue
            parts = line.split()
            if len(parts) >= 3 and parts[1] == workload and not line.startswith("#"):
                try:
                    out[parts[0]] = float(parts[2])
                except ValueError:
                    pass
    return out, commit


def build_cases(a):
    from pyugks_amd import cases
    if a.workload == "cavity_2x2":
        return [cases.cfg2_cavity_2x2(n_per_block=a.cells, nc=a.nc)], f"cavity_2x2x{a.cells}sq_nc{a.nc}_weno5"
    if a.workload == "cavity_64":
        return [cases.cfg1_cavity_gh()], "cavity_64sq_gh24_vanleer"
    if a.workload == "cavity_512":
        return [cases.cfg3_cavity_512(Kn=a.kn)], f"cavity_512sq_nc48_weno5_kn{a.kn:g}"
    if a.workload == "batch64":
        return cases.cfg4_batch(), "batch64_cavities_weno5"
    return [cases.cfg5_channel()], "channel_adsorbing_1024x256_nc40_weno5"


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


def run_precision(gs, precision, a, upd_per_step, workload):
    """Time one precision; returns its report and the final per-case state."""
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    sims = [UGKSim(copy.deepcopy(g)) for g in gs]
    mass0 = [sum(s.total_mass()) for s in sims]
    plan = StepPlan(sims, precision, graph=not a.no_graph)
    tuned = plan.autotune()
    print(f"[bench] {precision}: {len(plan.blocks)} blocks in {len(sims)} case(s), tuned {tuned}",
          file=sys.stderr, flush=True)
    wall_s, ev_s = timed_run(plan, a.steps, a.warmup)
    value = upd_per_step * a.steps / wall_s
    print(f"[bench] {precision}: {a.steps} steps in {wall_s:.4f} s ({ev_s:.4f} s by events)",
          file=sys.stderr, flush=True)
    kern = plan.run_timed(a.profile_steps)
    per_launch = {k: v / a.profile_steps for k, v in kern.items()}
    dom = max(per_launch, key=per_launch.get)
    b_alg = 32.0 if precision == "fp64" else 16.0
    step_s = wall_s / a.steps
    # whole step: the compulsory bytes of one step (1 read + 1 write of f per
    # update) over the measured time of the step (all eight launches)
    achieved = b_alg * upd_per_step / step_s / 1e9
    tag = f"{workload}_{precision}"
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "scope": "whole step (8 launches, graph replay)",
            "b_alg_per_update": b_alg, "updates_per_launch": upd_per_step}
    # the dominant kernel alone: the same algorithmic bytes over its mean
    # launch time (HIP events on the plan's stream)
    k_ach = b_alg * upd_per_step / (per_launch[dom] * 1e-3) / 1e9
    roof["dominant_kernel"] = {"kernel": dom, "ms_per_launch": per_launch[dom], "achieved": k_ach,
                               "frac": k_ach / HBM_PEAK_GBS}
    step_kernels = set(per_launch)          # the eight launches of a step (not the per-call refresh)
    traffic, t_commit = read_profile_table(f"hbm_traffic_per_kernel_{workload}.txt", tag)
    traffic = {k: v for k, v in traffic.items() if k in step_kernels}
    if traffic:
        roof["traffic"] = sum(traffic.values())
        roof["traffic_source"] = (f"profiles/hbm_traffic_per_kernel_{workload}.txt, rocprofv3 FETCH_SIZE + "
                                  f"WRITE_SIZE summed over the step's kernels, source commit {t_commit}")
        roof["traffic_over_alg"] = roof["traffic"] / (b_alg * upd_per_step)
    else:
        roof["traffic"] = None
    valu, v_commit = read_profile_table(f"valu_per_kernel_{workload}.txt", tag)
    valu = {k: v for k, v in valu.items() if k in step_kernels}
    if valu:
        per_step = sum(valu.values())
        roof["valu"] = {"insts_per_step": per_step,
                        "lane_ops_per_update": 64.0 * per_step / upd_per_step,
                        "frac": per_step / (step_s * CUS * CLOCK_HZ),
                        "dominant_kernel_frac": (valu.get(dom, 0.0) / (per_launch[dom] * 1e-3 * CUS * CLOCK_HZ)),
                        "peak": "one wave64 VALU instruction per CU per clock (256 CUs, 2.4 GHz)",
                        "source": f"profiles/valu_per_kernel_{workload}.txt, rocprofv3 SQ_INSTS_VALU, "
                                  f"source commit {v_commit}"}
    states = plan.states()
    final = [{"macro": [b.macro() for b in s.blocks], "q": [b.Q_D.download() for b in s.blocks],
              "mass": sum(s.total_mass())} for s in sims]
    rep = {"value": value, "ms_per_step": 1e3 * wall_s / a.steps, "roofline": roof,
           "kernel_ms_per_launch": per_launch, "flag": plan.flag(),
           "steps_done": [d.step for d in states], "autotune": tuned,
           "mass_drift": max(abs(f["mass"] - m0) / m0 for f, m0 in zip(final, mass0)),
           "graph": not a.no_graph}
    plan.close()
    for s in sims:
        for b in s.blocks:
            b.free_device()
    return rep, final


def per_kernel_path(gs, steps, upd_per_step):
    """The reference's call sequence (UGKSim.one_step: one launch per reference
    kernel and block, with its host readbacks of the flag and time step), in
    FP64, case by case: 2 untimed steps, then `steps` timed ones; the times of
    the cases add up to one step of the workload."""
    from pyugks_amd import lib as L
    from pyugks_amd.sim import UGKSim
    total = 0.0
    for g in gs:
        sim = UGKSim(copy.deepcopy(g))
        for _ in range(2):
            sim.one_step(True)
        L.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            sim.one_step(True)
        L.synchronize()
        total += (time.perf_counter() - t0) / steps
        for b in sim.blocks:
            b.free_device()
    print(f"[bench] per-kernel path: {1e3 * total:.3f} ms per step", file=sys.stderr, flush=True)
    return {"ms_per_step": 1e3 * total, "value": upd_per_step / total, "unit": "cell-velocity updates/s",
            "steps": steps, "dtype": "f64", "path": "per-kernel (UGKSim.one_step)"}


def errors_between(final_a, final_b):
    from pyugks_amd.metrics import field_errors
    worst = {}
    for fa, fb in zip(final_a, final_b):
        e = field_errors(fa["macro"], fb["macro"], fa["q"], fb["q"])
        for k, v in e.items():
            worst[k] = max(worst.get(k, 0.0), v)
    return worst


def case_updates(g):
    gp = copy.deepcopy(g).prepare()
    return sum(b.ni * b.nj for b in gp.blocks) * gp.Nv


def cpu_sample(gs, workload):
    """The bounded CPU sample of a workload: the whole case for the single-case
    workloads; for the 64-case batch, four cases spread over its sizes (ranks
    15, 31, 47 and 63 by updates per step, the largest included; the oracle
    holds four FP64 f-sized arrays per block)."""
    if workload.startswith("batch") and len(gs) > 4:
        order = sorted(range(len(gs)), key=lambda i: case_updates(gs[i]))
        idx = sorted(order[len(gs) // 4 - 1::len(gs) // 4][:4])
        return [gs[i] for i in idx], f"cases {idx} of the {len(gs)}-case batch (four sizes from its range)"
    return gs, "the whole workload"


# A WENO5 cavity at CFL 0.8 ends in NaN after 75 to 160 steps (the scheme as the
# reference writes it); the CPU sample stays well inside that horizon.
CPU_MAX_STEPS_PER_RUN = 10


def cpu_baseline(gs, a, workload):
    """The FP64 oracle on the host CPU over a bounded sample of the workload:
    all threads of this process's CPU share, then one thread, each the median
    of three timed runs; then the FP64 GPU step over the same steps from the
    same initial state, compared with the oracle's state."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    sub, what = cpu_sample(gs, workload)
    upd = sum(case_updates(g) for g in sub)
    cores, nproc = host_cores()
    oras = [O.OracleSim(copy.deepcopy(g), threads=cores) for g in sub]
    steps_done = 0

    def run(k):
        nonlocal steps_done
        t0 = time.perf_counter()
        for o in oras:
            o.run(k)
        steps_done += k
        return time.perf_counter() - t0

    def median_rate(threads):
        O.set_threads(threads)
        one = run(1)                                   # untimed: sizes the runs
        k = max(1, min(CPU_MAX_STEPS_PER_RUN, int(round(a.cpu_seconds / max(one, 1e-3)))))
        rates = []
        for r in range(3):
            dt = run(k)
            rates.append(upd * k / dt)
            print(f"[bench] cpu baseline, {threads} thread(s): run {r + 1}/3, {k} step(s) in {dt:.1f} s",
                  file=sys.stderr, flush=True)
        return sorted(rates)[1], rates, k

    print(f"[bench] cpu baseline on {what}: {upd:.3e} updates per step, {cores} threads of {nproc}",
          file=sys.stderr, flush=True)
    v_all, r_all, k_all = median_rate(cores)
    v_one, r_one, k_one = median_rate(1)
    out = {"value": v_all, "unit": "cell-velocity updates/s", "cores": cores, "kind": "port",
           "sample": f"{what}; FP64 C/OpenMP oracle; median of 3 runs of {k_all} step(s) on {cores} threads "
                     f"(the CPU share of this process; the machine has {nproc}); the 1-thread figure is the "
                     f"median of 3 runs of {k_one} step(s)",
           "runs": r_all, "one_core": {"value": v_one, "runs": r_one}, "nproc": nproc, "cpu_model": cpu_model()}
    sims = [UGKSim(copy.deepcopy(g)) for g in sub]
    plan = StepPlan(sims, "fp64")
    plan.run(steps_done)
    final_gpu = [{"macro": [b.macro() for b in s.blocks], "q": [b.Q_D.download() for b in s.blocks]} for s in sims]
    final_ora = [{"macro": [b["macro"] for b in o.blocks], "q": [b["Q"] for b in o.blocks]} for o in oras]
    out["fp64_gpu_vs_oracle"] = errors_between(final_gpu, final_ora)
    out["fp64_gpu_vs_oracle"]["steps"] = steps_done
    plan.close()
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: runs on one GPU only; --gpus {a.gpus} is not supported (blocks would be "
              "sharded one group per GPU with RCCL ghost exchange, which is not built)", file=sys.stderr)
        sys.exit(2)
    from pyugks_amd import lib as L
    L.load()
    if L.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return

    gs, workload = build_cases(a)
    print(f"[bench] init {workload}", file=sys.stderr, flush=True)
    upd_per_step = 0
    nv_set = set()
    for g in gs:
        gp = copy.deepcopy(g).prepare()
        for b in gp.blocks:
            upd_per_step += b.ni * b.nj * gp.Nv
        nv_set.add(gp.Nv)
    prec = ["fp64", "mixed"] if a.precision == "both" else [a.precision]
    reps, finals = {}, {}
    for p in prec:
        reps[p], finals[p] = run_precision(gs, p, a, upd_per_step, workload)
    head = "mixed" if "mixed" in reps else prec[0]
    r = reps[head]
    out = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "cell-velocity updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32 storage/reconstruction/flux, f64 moments/update/dt",
        "data": "synthetic (Maxwellian initial state, seeded perturbations; no input files)",
        "config": {"workload": workload, "cases": len(gs), "updates_per_step": upd_per_step,
                   "Nv": sorted(nv_set), "path": f"fused {head}" + (" graph" if r["graph"] else "")},
        "roofline": r["roofline"],
        "kernel_ms_per_launch": r["kernel_ms_per_launch"],
        "autotune": r["autotune"],
        "mass_drift": r["mass_drift"],
        "flag": r["flag"],
    }
    if "fp64" in reps and "mixed" in reps:
        f = reps["fp64"]
        assert f["steps_done"] == r["steps_done"]
        out["speedup_vs_fp64"] = r["value"] / f["value"]
        out["err_vs_fp64"] = errors_between(finals["mixed"], finals["fp64"])
        out["fp64"] = {"value": f["value"], "ms_per_step": f["ms_per_step"], "dtype": "f64",
                       "roofline": f["roofline"], "mass_drift": f["mass_drift"],
                       "kernel_ms_per_launch": f["kernel_ms_per_launch"], "flag": f["flag"]}
    finals.clear()
    if a.per_kernel_steps > 0:
        out["per_kernel_path"] = per_kernel_path(gs, a.per_kernel_steps, upd_per_step)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(gs, a, workload)
        out["cpu_baseline"]["gpu_over_cpu"] = r["value"] / out["cpu_baseline"]["value"]
    if any(rep["flag"] for rep in reps.values()):
        # e.g. a WENO5 cavity at CFL 0.8 run past about 75 steps: the scheme
        # itself ends in NaN there; the step times stay valid, the errors do not
        print("[bench] warning: a run raised the reference's error flag (NaN); its error and drift "
              "fields are reported as null", file=sys.stderr, flush=True)
    print(json.dumps(finite_json(out), allow_nan=False), flush=True)


def finite_json(x):
    """x with every NaN or infinity replaced by None, so the line stays strict JSON."""
    if isinstance(x, dict):
        return {k: finite_json(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [finite_json(v) for v in x]
    if isinstance(x, float) and not math.isfinite(x):
        return None
    return x


if __name__ == "__main__":
    main()
