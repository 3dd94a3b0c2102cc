"""Long-horizon accuracy of the mixed-precision step against the FP64 step.

For the 512^2 cavity at each of its Knudsen numbers (Newton-Cotes 49^2,
WENO5, rho perturbed by 1%), for the adsorbing-wall channel (8 blocks,
1024 x 256, 41^2) and for the 64-case batch of cavities (one plan, each
case on its own time step), both fused plans run the same cases from the
same initial state; at each checkpoint the relative L-inf and L2
differences of rho, U, T and q between them and the mass drift of each are
printed, one JSON line per checkpoint (for the batch: the worst case of
each figure, every case measured on its own scale).  The FP64 plan follows the FP64 oracle to 1e-12 per step
(the parity tests), so it stands in for the oracle over horizons the oracle
cannot reach in reasonable time.  A run whose NaN flag trips reports it and
stops at that checkpoint.

    python -m pyugks_amd.accuracy [--cfl 0.4] [--steps 50,100,200,500]
                                  [--cases cavity_512_kn0.001,...] [--method WENO5]
"""
from __future__ import annotations

import argparse
import copy
import json
import sys

import numpy as np

from . import cases
from .metrics import field_errors, mass_drift

CASES = {
    "cavity_512_kn0.001": lambda: cases.cfg3_cavity_512(Kn=0.001),
    "cavity_512_kn0.075": lambda: cases.cfg3_cavity_512(Kn=0.075),
    "cavity_512_kn1": lambda: cases.cfg3_cavity_512(Kn=1.0),
    "channel": lambda: cases.cfg5_channel(),
    "batch64": lambda: cases.cfg4_batch(),
}


def _run(gs, precision, checkpoints):
    """Snapshots (per case: macro, q, mass drift; the plan's flag; the first
    case's time) of one plan over the cases gs at each checkpoint."""
    from .fused import StepPlan
    from .sim import UGKSim
    sims = [UGKSim(copy.deepcopy(g)) for g in gs]
    m0 = [sum(s.total_mass()) for s in sims]
    plan = StepPlan(sims, precision, graph=True)
    out, done = [], 0
    for k in checkpoints:
        plan.run(k - done)
        done = k
        plan.states()
        flag = plan.flag()
        t = sims[0].gdata.time
        out.append({"macro": [[b.macro() for b in s.blocks] for s in sims],
                    "q": [[b.Q_D.download() for b in s.blocks] for s in sims],
                    "drift": [mass_drift(a, sum(s.total_mass())) for s, a in zip(sims, m0)],
                    "flag": flag, "time": t})
        print(f"[accuracy] {precision}: step {k}, t = {t:.6g}, flag {flag:g}", file=sys.stderr, flush=True)
        if flag != 0.0:
            break
    plan.close()
    for s in sims:
        for b in s.blocks:
            b.free_device()    # tens of GB at these sizes: released now, not at garbage collection
    return out


def _worst(mix, ref):
    """field_errors of every case, the worst of each figure over the cases;
    with several cases also the median case's U error and how many cases
    exceed 1e-5 in rho, U or T (L-inf)."""
    worst, u, over = {}, [], 0
    for i in range(len(ref["macro"])):
        e = field_errors(mix["macro"][i], ref["macro"][i], mix["q"][i], ref["q"][i])
        for key, v in e.items():
            worst[key] = max(worst.get(key, 0.0), v)
        u.append(e["U_linf"])
        over += max(e["rho_linf"], e["U_linf"], e["T_linf"]) > 1e-5
    if len(u) > 1:
        worst["U_linf_median_case"] = float(sorted(u)[len(u) // 2])
        worst["cases_over_1e-5"] = over
    return worst


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--cfl", type=float, default=0.4)
    ap.add_argument("--steps", default="50,100,200,500")
    ap.add_argument("--cases", default=",".join(CASES))
    ap.add_argument("--method", default=None, help="reconstruction (default: the case's own, WENO5)")
    a = ap.parse_args(argv)
    checkpoints = sorted(int(s) for s in a.steps.split(","))
    for name in a.cases.split(","):
        gs = CASES[name]()
        gs = gs if isinstance(gs, list) else [gs]
        for g in gs:
            g.CFL = a.cfl
            if a.method:
                g.flux_method = a.method
        ref = _run(gs, "fp64", checkpoints)
        mix = _run(gs, "mixed", checkpoints)
        for k, r, m in zip(checkpoints, ref, mix):
            line = {"case": name, "cases": len(gs), "cfl": a.cfl, "method": gs[0].flux_method, "step": k,
                    "time": r["time"], "flag_fp64": r["flag"], "flag_mixed": m["flag"],
                    "drift_fp64": max(r["drift"]), "drift_mixed": max(m["drift"])}
            if r["flag"] == 0.0 and m["flag"] == 0.0:
                line.update(_worst(m, r))
            print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
