"""The host step loop (ugks_sim.py:503-678) on the GPU: the last step is
shortened onto max_time, the residual monitor and its stop criteria, the
open-domain mass monitor, and the same loop driven through a fused plan."""
import copy

import numpy as np
import pytest

from helpers import rel_err
from pyugks_amd import cases

pytestmark = pytest.mark.gpu


def _cavity():
    return cases.cavity(8, ("NC", 8), "WENO5", 0.05, lid=0.2, n_cells_j=6)


@pytest.mark.parametrize("fused", [False, True])
def test_run_lands_on_max_time(hip, fused):
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    print(f"[gpu] host loop max_time fused={fused}", flush=True)
    sim = UGKSim(copy.deepcopy(_cavity()))
    t_end = 5.5 * sim.gdata.dt
    plan = StepPlan(sim, "fp64") if fused else None
    assert sim.run(nsteps=100, max_time=t_end, plan=plan, log=None)
    assert sim.exit_reason == "time limit reached"
    assert sim.force_time
    assert abs(sim.gdata.time - t_end) <= 1e-15 * t_end
    assert sim.step in (6, 7)


def test_fused_and_per_kernel_loops_agree(hip):
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    a, b = UGKSim(copy.deepcopy(_cavity())), UGKSim(copy.deepcopy(_cavity()))
    t_end = 4.25 * a.gdata.dt
    a.run(nsteps=50, max_time=t_end, log=None)
    b.run(nsteps=50, max_time=t_end, plan=StepPlan(b, "fp64"), log=None)
    assert a.step == b.step and abs(a.gdata.time - b.gdata.time) <= 1e-14 * t_end
    assert rel_err(b.blocks[0].macro(), a.blocks[0].macro()) <= 1e-12
    assert rel_err(b.blocks[0].f(), a.blocks[0].f()) <= 1e-12


@pytest.mark.parametrize("fused", [False, True])
def test_residual_monitor_and_minimum_residual_stop(hip, fused):
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    sim = UGKSim(copy.deepcopy(_cavity()))
    r = sim.gdata.residual_options
    r.get_residual, r.residual_count, r.residual_start, r.slope_start = True, 2, 6, 100
    r.min_residual = 1e-300
    assert sim.run(nsteps=8, plan=StepPlan(sim, "fp64") if fused else None, log=None)
    assert sim.exit_reason == "step count reached"
    assert sim.time_history_residual_N == [2, 4, 6, 8]      # recorded after the step (step + 1) % 2 == 0
    hist = np.array(sim.time_history_residual)
    assert hist.shape == (4, 4) and np.all(hist > 0) and np.all(np.isfinite(hist))
    # every normalised residual is far below 1: the next run stops on the first check
    r.min_residual = 1.0
    sim.run(nsteps=20, plan=StepPlan(sim, "fp64") if fused else None, log=None)
    assert sim.exit_reason == "minimum residual reached"


def test_open_domain_mass_monitor(hip):
    from pyugks_amd.sim import UGKSim
    g = cases.cfg5_channel(widths=[4, 4], height=4, nc=8)
    sim = UGKSim(copy.deepcopy(g))
    assert sim.is_open
    r = sim.gdata.residual_options
    r.get_residual, r.residual_count, r.slope_start, r.residual_start = True, 1, 100, 100
    sim.run(nsteps=3, log=None)
    res = sim.gdata.residual_options.global_residual
    assert res.shape == (6,)
    assert res[4] > 0                       # |boundary mass flow|: the inflow carries mass in
    bulk, ads = sim.total_mass()
    assert abs(res[5] - abs(sim.initial_mass_total - (bulk + ads)) / sim.initial_mass_total) < 1e-14


@pytest.mark.parametrize("fused", [False, True])
def test_growth_limit_binding_in_the_flux_loop(hip, fused):
    """dt set to a third of what the blocks allow: the first step computes the
    flux with dt, re-evaluates (the growth limit doubles it), computes the
    flux again with 2/3 of the allowance, re-evaluates once more and updates
    with the full allowance (ugks_sim.py:430-446); later steps grow dt to the
    limit in turn.  Both the per-kernel loop and the fused plan follow the
    oracle step for step: dt, time, f and macro within 1e-12."""
    from oracle import OracleSim
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    g = cases.cavity(8, ("NC", 8), "WENO5", 0.05, lid=0.2, n_cells_j=6, rho_perturb=0.01)
    sim = UGKSim(copy.deepcopy(g))
    ora = OracleSim(copy.deepcopy(g), threads=4)
    M = sim.gdata.dt
    assert abs(ora.dt - M) <= 1e-15 * M
    sim.gdata.dt = M / 3.0
    ora.sim.dt = M / 3.0
    plan = StepPlan(sim, "fp64") if fused else None
    print(f"[gpu] growth limit in the flux loop, fused={fused}", flush=True)
    for step in range(4):
        if fused:
            plan.run(1)
            d = plan.state()
            dt, t = d.dt, d.time
        else:
            _, err = sim.one_step(True)
            assert not err
            dt, t = sim.gdata.dt, sim.gdata.time
        ora.one_step(True)
        assert abs(dt - ora.dt) <= 1e-12 * ora.dt, (step, dt, ora.dt)
        assert abs(t - ora.time) <= 1e-12 * ora.time, (step, t, ora.time)
        b, o = sim.blocks[0], ora.blocks[0]
        assert rel_err(b.macro(), o["macro"]) <= 1e-12, step
        assert rel_err(b.f(), o["f"]) <= 1e-12, step
