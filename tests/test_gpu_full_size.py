"""The benchmark configurations at their full sizes.

- The 64^2 Gauss-Hermite vanLeer cavity (the reference's own CPU case):
  twenty FP64 steps against the oracle at 1e-12 per step, and the mixed
  plan within 1e-5 of it.
- The 2x2x128^2 cavity with 33^2 velocities through the per-kernel path:
  three steps against the oracle at 1e-12 per step.
- The 64-case batch, the 512^2 cavity and the adsorbing channel are too
  large for the oracle to follow in a test, so they are checked through
  properties that do not depend on size: the mixed plan within 1e-5 (rho,
  U, T, L-inf relative) of the FP64 plan, which follows the oracle to
  1e-12 per step on the smaller cases; mass conserved in the closed
  cavities (to rounding in FP64, to the FP32 storage of f in mixed); the
  cases of the batch on their own time steps; no error flag.
"""
import copy
import gc
import os

import numpy as np
import pytest

from helpers import SCALE_FLOOR, field_errors, rel_err
from pyugks_amd import cases

TOL = 1e-12
pytestmark = pytest.mark.gpu


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def _run(gs, precision, nsteps):
    """Step the cases in one graph-replayed plan; their macro fields, heat
    fluxes, mass drifts and time steps, and the plan's flag, after nsteps."""
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    sims = [UGKSim(copy.deepcopy(g)) for g in gs]
    m0 = [sum(s.total_mass()) for s in sims]
    plan = StepPlan(sims, precision, graph=True)
    plan.run(nsteps)
    states = plan.states()
    out = {"macro": [[b.macro_D.download() for b in s.blocks] for s in sims],
           "q": [[b.Q_D.download() for b in s.blocks] for s in sims],
           "drift": [abs(sum(s.total_mass()) - a) / a for s, a in zip(sims, m0)],
           "dt": [d.dt for d in states], "flag": plan.flag()}
    plan.close()
    # the simulations hold tens of GB of device memory in reference cycles:
    # release it now, not whenever the cyclic collector next runs
    for sim in sims:
        for b in sim.blocks:
            b.free_device()
    del plan, sims
    gc.collect()
    return out


def test_cavity_64_gh24_vanleer_fp64_and_mixed_against_the_oracle(hip):
    from oracle import OracleSim
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    g = cases.cfg1_cavity_gh()
    print("[gpu] full size: 64^2 cavity, Gauss-Hermite 24^2, vanLeer, 20 steps", flush=True)
    ora = OracleSim(copy.deepcopy(g), threads=_threads())
    gpu = UGKSim(copy.deepcopy(g))
    plan = StepPlan(gpu, "fp64", graph=True)
    for step in range(20):
        plan.run(1)
        ora.one_step(True)
        d = plan.state()
        assert abs(d.dt - ora.dt) <= TOL * ora.dt, step
        bg, bo = gpu.blocks[0], ora.blocks[0]
        assert rel_err(bg.macro_D.download(), bo["macro"]) <= TOL, step
        assert rel_err(bg.f_D.download(), bo["f"]) <= TOL, step
    assert rel_err(gpu.blocks[0].Q_D.download(), ora.blocks[0]["Q"], SCALE_FLOOR["Q"]) <= TOL
    plan.close()
    gpu.blocks[0].free_device()
    mixed = _run([g], "mixed", 20)
    e = field_errors(mixed["macro"][0], [ora.blocks[0]["macro"]])
    print(f"[gpu] mixed vs oracle after 20 steps: {e['rho_linf']:.2e} {e['U_linf']:.2e} {e['T_linf']:.2e}",
          flush=True)
    for k in ("rho_linf", "U_linf", "T_linf"):
        assert e[k] <= 1e-5, (k, e[k])
    assert mixed["flag"] == 0.0


def test_cavity_2x2x128_per_kernel_path_against_the_oracle(hip):
    """The reference-shaped per-kernel path (UGKSim.one_step) at the 2x2x128^2
    size with 33^2 velocities: four blocks of 128^2 cells, 1089 velocities
    (17 full wavefront rounds and one lane of an 18th per face or cell), the
    oracle's f, macro fields and heat flux to 1e-12 after each of 3 steps."""
    from oracle import OracleSim
    from pyugks_amd.sim import UGKSim
    g = cases.cfg2_cavity_2x2()
    print("[gpu] full size: 2x2x128^2 cavity, 33^2 velocities, per-kernel path, 3 steps", flush=True)
    ora = OracleSim(copy.deepcopy(g), threads=_threads())
    gpu = UGKSim(copy.deepcopy(g))
    for step in range(3):
        gpu.one_step(True)
        ora.one_step(True)
        assert abs(gpu.gdata.dt - ora.dt) <= TOL * ora.dt, step
        for bg, bo in zip(gpu.blocks, ora.blocks):
            assert rel_err(bg.macro_D.download(), bo["macro"]) <= TOL, step
            assert rel_err(bg.f_D.download(), bo["f"]) <= TOL, step
    for bg, bo in zip(gpu.blocks, ora.blocks):
        assert rel_err(bg.Q_D.download(), bo["Q"], SCALE_FLOOR["Q"]) <= TOL
        bg.free_device()
    del gpu, ora
    gc.collect()


def test_batch_of_64_cases_mixed_tracks_fp64_and_conserves_mass(hip):
    gs = cases.cfg4_batch()
    print("[gpu] full size: 64-case batch, 10 steps in each precision", flush=True)
    ref = _run(gs, "fp64", 10)
    mix = _run(gs, "mixed", 10)
    assert ref["flag"] == 0.0 and mix["flag"] == 0.0
    # the cases step with their own time steps (cases of one size and velocity
    # grid share one), and the two precisions agree on each
    assert len({round(x, 12) for x in ref["dt"]}) > len(gs) // 4
    assert np.allclose(mix["dt"], ref["dt"], rtol=1e-6, atol=0)
    worst = {}
    for i in range(len(gs)):
        e = field_errors(mix["macro"][i], ref["macro"][i], mix["q"][i], ref["q"][i])
        for k in ("rho_linf", "U_linf", "T_linf"):
            worst[k] = max(worst.get(k, 0.0), e[k])
            assert e[k] <= 1e-5, (i, k, e[k])
    print(f"[gpu] worst case of 64: {worst}; mass drift FP64 {max(ref['drift']):.1e}, mixed {max(mix['drift']):.1e}",
          flush=True)
    assert max(ref["drift"]) <= 1e-13
    assert max(mix["drift"]) <= 1e-8


@pytest.mark.parametrize("kn", [0.001, 1.0])
def test_cavity_512_mixed_tracks_fp64_and_conserves_mass(hip, kn):
    g = cases.cfg3_cavity_512(Kn=kn)
    print(f"[gpu] full size: 512^2 cavity, 49^2 velocities, Kn {kn}, 5 steps in each precision", flush=True)
    ref = _run([g], "fp64", 5)
    mix = _run([g], "mixed", 5)
    assert ref["flag"] == 0.0 and mix["flag"] == 0.0
    e = field_errors(mix["macro"][0], ref["macro"][0])
    for k in ("rho_linf", "U_linf", "T_linf"):
        assert e[k] <= 1e-5, (k, e[k])
    assert ref["drift"][0] <= 1e-13 and mix["drift"][0] <= 1e-9


def test_adsorbing_channel_mixed_tracks_fp64(hip):
    g = cases.cfg5_channel()
    print("[gpu] full size: adsorbing channel, 8 blocks, 1024x256, 41^2 velocities, 5 steps", flush=True)
    ref = _run([g], "fp64", 5)
    mix = _run([g], "mixed", 5)
    assert ref["flag"] == 0.0 and mix["flag"] == 0.0
    e = field_errors(mix["macro"][0], ref["macro"][0])
    for k in ("rho_linf", "U_linf", "T_linf"):
        assert e[k] <= 1e-5, (k, e[k])
    # the open channel's mass changes (inflow, outflow, adsorption), the same in both
    assert abs(mix["drift"][0] - ref["drift"][0]) <= 1e-5 * max(ref["drift"][0], 1e-12) + 1e-9


def test_accuracy_driver_reports_each_checkpoint(hip, capsys):
    """python -m pyugks_amd.accuracy: one JSON line per checkpoint with the
    mixed-vs-FP64 errors and both mass drifts (two short checkpoints of the
    channel here)."""
    import json

    from pyugks_amd import accuracy
    print("[gpu] accuracy driver, channel, 2 checkpoints", flush=True)
    capsys.readouterr()
    accuracy.main(["--cases", "channel", "--steps", "1,2", "--cfl", "0.4"])
    lines = [json.loads(x) for x in capsys.readouterr().out.splitlines() if x.startswith("{")]
    assert [d["step"] for d in lines] == [1, 2]
    for d in lines:
        assert d["case"] == "channel" and d["flag_fp64"] == 0.0 and d["flag_mixed"] == 0.0
        assert d["U_linf"] <= 1e-5 and d["rho_linf"] <= 1e-5 and d["T_linf"] <= 1e-5
        assert abs(d["drift_mixed"] - d["drift_fp64"]) <= 1e-9
