"""Long runs of the fused FP64 step against the FP64 oracle.

The WENO5 variant of the scheme (clUGKS.c:110-199 with the WENO5 of
clFunctions.c:268-312, then the single-stage UGKS update) is not stable at
CFL 0.8 on the lid-driven cavity: a few interior cells heat up and the run
ends in NaN after about a hundred steps, in the oracle as on the GPU.  At
CFL 0.4 the same case runs on.  These tests check that the GPU reproduces
both: the long stable run to the per-step bound (1e-12) after 300 steps, and
the failing run to the same step, reported through the reference's flag
protocol (clError.c: the NaN flag set by UGKS_update).
"""
import copy

import numpy as np
import pytest

from helpers import rel_err
from pyugks_amd import cases

pytestmark = pytest.mark.gpu


def weno_cavity(cfl):
    return cases.cavity(32, ("NC", 32), "WENO5", 0.01, CFL=cfl)


def test_weno5_cavity_at_cfl_0p4_matches_oracle_after_300_steps(hip):
    from oracle import OracleSim
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    print("[gpu] long run: 32^2 WENO5 cavity, NC 33^2, CFL 0.4, 300 steps", flush=True)
    g = weno_cavity(0.4)
    gpu = UGKSim(copy.deepcopy(g))
    plan = StepPlan(gpu, "fp64")
    ora = OracleSim(copy.deepcopy(g), threads=8)
    plan.run(300)
    ora.run(300)
    d = plan.state()
    assert d.step == ora.sim.step == 300
    assert abs(d.time - ora.time) <= 1e-12 * ora.time
    assert plan.flag() == 0.0
    bg, bo = gpu.blocks[0], ora.blocks[0]
    ef, em = rel_err(bg.f_D.download(), bo["f"]), rel_err(bg.macro_D.download(), bo["macro"])
    print(f"[gpu] after 300 steps: f {ef:.2e}, macro {em:.2e}", flush=True)
    assert ef <= 1e-12 and em <= 1e-12, (ef, em)     # 1.6e-14 and 7.7e-15 measured
    m = bo["macro"]
    assert np.isfinite(m).all() and m[..., 0].min() > 0.8 and m[..., 3].max() < 1.1
    plan.close()


def test_weno5_cavity_at_cfl_0p8_flags_nan_at_the_oracles_step(hip):
    from oracle import OracleSim
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    print("[gpu] failing run: 32^2 WENO5 cavity, NC 33^2, CFL 0.8, until the NaN flag", flush=True)
    g = weno_cavity(0.8)
    ora = OracleSim(copy.deepcopy(g), threads=8)
    fail = None
    for k in range(1, 201):
        try:
            ora.one_step(True)
        except RuntimeError:
            fail = k
            break
    assert fail is not None and fail > 100, fail
    gpu = UGKSim(copy.deepcopy(g))
    plan = StepPlan(gpu, "fp64")
    plan.run(fail - 1)
    assert plan.flag() == 0.0
    plan.run(1)
    print(f"[gpu] oracle failed at step {fail}; GPU flag after it: {plan.flag()}", flush=True)
    assert plan.flag() != 0.0
    plan.close()
