"""The allocation-free step (ugks_step_run_noalloc) against the plan's step.

Two calls in a row on the same GPU, the second with a smaller case (fewer
blocks, cells and velocities, another reconstruction), each bit for bit the
plan's result from the same initial state: the library's static working
memory keeps the first call's contents, and the second call must not read
any of them.  A case beyond the working memory is refused before anything
runs.
"""
import copy

import numpy as np
import pytest

from pyugks_amd import cases

pytestmark = pytest.mark.gpu


def _same(a, b, plan=None):
    for x, y in zip(a.blocks, b.blocks):
        fx = plan.f64(x) if plan is not None else x.f()
        assert np.array_equal(fx, y.f())
        assert np.array_equal(x.macro(), y.macro())
        assert np.array_equal(x.Q_D.download(), y.Q_D.download())
        assert np.array_equal(x.wall_fluxes_D.download(), y.wall_fluxes_D.download())
        assert np.array_equal(x.flag_D.download(), y.flag_D.download())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_step_equals_the_plan_step_on_two_calls(hip, precision):
    from pyugks_amd.fused import StepPlan, run_noalloc
    from pyugks_amd.sim import UGKSim
    big = cases.cavity(16, ("NC", 12), "WENO5", 0.02, nb=(2, 1), n_cells_j=12, rho_perturb=0.01)
    small = cases.cavity(8, ("NC", 8), "vanLeer", 0.05, lid=0.2, n_cells_j=6)
    for g, nsteps in ((big, 3), (small, 2)):
        print(f"[gpu] noalloc {precision}: {len(g.blocks)} block(s), {nsteps} steps", flush=True)
        a, b = UGKSim(copy.deepcopy(g)), UGKSim(copy.deepcopy(g))
        plan = StepPlan(a, precision)
        plan.run(nsteps)
        d = plan.state()
        e = run_noalloc(b, precision, nsteps)
        assert (b.gdata.dt, b.gdata.time, b.step) == (d.dt, d.time, d.step)
        assert e.max_dt > 0 and e.overshoot == 0
        _same(a, b, plan if precision == "mixed" else None)
        plan.close()


def test_noalloc_refuses_a_case_beyond_its_working_memory(hip):
    from pyugks_amd import lib as L
    from pyugks_amd.fused import run_noalloc
    from pyugks_amd.sim import UGKSim
    cap = L.load().ugks_step_noalloc_capacity()
    assert cap == 64 << 20
    g = cases.cavity(280, ("NC", 4), "vanLeer", 0.05, n_cells_j=280)      # ~157k faces: about 71 MB
    sim = UGKSim(g)
    f0 = sim.blocks[0].f()
    with pytest.raises(L.UGKSError, match="working memory"):
        run_noalloc(sim, "fp64", 1)
    assert np.array_equal(sim.blocks[0].f(), f0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_step_with_time_dependent_walls_and_joins(hip, precision):
    """A case whose lid speed depends on time (a wall-value expression the
    step re-evaluates every step) over two joined blocks, on an adsorbing
    channel's walls: the allocation-free step equals the plan's, bit for bit."""
    from pyugks_amd import bc as bcd
    from pyugks_amd.fused import StepPlan, run_noalloc
    from pyugks_amd.sim import UGKSim
    g = cases.cavity(10, ("NC", 8), "WENO5", 0.05, lid=0.15, n_cells_j=8, nb=(2, 1))
    for blk in g.blocks:
        blk.bc_list[bcd.NORTH] = bcd.DiffuseBC(D=1.0, T=1.0, U=0.15, UDF_U="0.2*sin(2*PI*t/0.05) + 0.05*s")
    ch = cases.cfg5_channel(widths=[6, 8], height=6, nc=8, Kn=0.05)
    for gg, nsteps in ((g, 3), (ch, 2)):
        print(f"[gpu] noalloc {precision}: {len(gg.blocks)} blocks, {nsteps} steps", flush=True)
        a, b = UGKSim(copy.deepcopy(gg)), UGKSim(copy.deepcopy(gg))
        plan = StepPlan(a, precision)
        plan.run(nsteps)
        d = plan.state()
        run_noalloc(b, precision, nsteps)
        assert (b.gdata.dt, b.gdata.time, b.step) == (d.dt, d.time, d.step)
        _same(a, b, plan if precision == "mixed" else None)
        for x, y in zip(a.blocks, b.blocks):
            assert np.array_equal(x.wall_prop_D.download(), y.wall_prop_D.download())
            assert np.array_equal(x.wall_cover_D.download(), y.wall_cover_D.download())
        plan.close()
