"""Long-horizon accuracy of the mixed-precision step against the FP64 step.

For the 512^2 cavity at each of its Knudsen numbers (Newton-Cotes 49^2,
WENO5, rho perturbed by 1%) and for the adsorbing-wall channel (8 blocks,
1024 x 256, 41^2), both fused plans run the same case from the same initial
state; at each checkpoint the relative L-inf and L2 differences of rho, U,
T and q between them and the mass drift of each are printed, one JSON line
per checkpoint.  The FP64 plan follows the FP64 oracle to 1e-12 per step
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
}


def _run(g, precision, checkpoints):
    """Snapshots (macro, q, mass, flag, time) of one plan at each checkpoint."""
    from .fused import StepPlan
    from .sim import UGKSim
    sim = UGKSim(copy.deepcopy(g))
    m0 = sum(sim.total_mass())
    plan = StepPlan(sim, precision, graph=True)
    out, done = [], 0
    for k in checkpoints:
        plan.run(k - done)
        done = k
        plan.states()
        flag = plan.flag()
        out.append({"macro": [b.macro() for b in sim.blocks], "q": [b.Q_D.download() for b in sim.blocks],
                    "drift": mass_drift(m0, sum(sim.total_mass())), "flag": flag, "time": sim.gdata.time})
        print(f"[accuracy] {precision}: step {k}, t = {sim.gdata.time:.6g}, flag {flag:g}", file=sys.stderr,
              flush=True)
        if flag != 0.0:
            break
    plan.close()
    for b in sim.blocks:
        b.free_device()        # tens of GB at these sizes: released now, not at garbage collection
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--cfl", type=float, default=0.4)
    ap.add_argument("--steps", default="50,100,200,500")
    ap.add_argument("--cases", default=",".join(CASES))
    ap.add_argument("--method", default=None, help="reconstruction (default: the case's own, WENO5)")
    a = ap.parse_args(argv)
    checkpoints = sorted(int(s) for s in a.steps.split(","))
    for name in a.cases.split(","):
        g = CASES[name]()
        g.CFL = a.cfl
        if a.method:
            g.flux_method = a.method
        ref = _run(g, "fp64", checkpoints)
        mix = _run(g, "mixed", checkpoints)
        for k, r, m in zip(checkpoints, ref, mix):
            line = {"case": name, "cfl": a.cfl, "method": g.flux_method, "step": k, "time": r["time"],
                    "flag_fp64": r["flag"], "flag_mixed": m["flag"],
                    "drift_fp64": r["drift"], "drift_mixed": m["drift"]}
            if r["flag"] == 0.0 and m["flag"] == 0.0:
                line.update(field_errors(m["macro"], r["macro"], m["q"], r["q"]))
            print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
