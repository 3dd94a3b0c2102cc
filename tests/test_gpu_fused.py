"""GPU parity of the fused, batched step against the FP64 oracle.

FP64 plan: f, macro, Q, flux_macro, wall state, dt and time after every
step within 1e-12 relative (max|a-b| / max|b|; heat fluxes against the
unit-state scale).  Mixed FP32/FP64 plan: L-inf relative error of rho, U,
T below 1e-5 after the steps run.
"""
import copy

import numpy as np
import pytest

from helpers import SCALE_FLOOR, field_errors, rel_err
from pyugks_amd import cases
from test_gpu_reference_path import CASES

TOL = 1e-12
pytestmark = pytest.mark.gpu


def build(name, precision):
    from oracle import OracleSim
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    print(f"[gpu] fused {precision} case {name}", flush=True)
    gpu = UGKSim(copy.deepcopy(CASES[name]))
    plan = StepPlan(gpu, precision)
    ora = OracleSim(copy.deepcopy(CASES[name]), threads=4)
    return gpu, plan, ora


def interior_faces(arr, G, ni, nj):
    return arr[G:G + ni + 1, G:G + nj + 1]


@pytest.mark.parametrize("name", list(CASES))
def test_fused_fp64_matches_oracle(hip, name):
    gpu, plan, ora = build(name, "fp64")
    for step in range(4):
        plan.run(1)
        ora.one_step(True)
        d = plan.state()
        assert abs(d.dt - ora.dt) <= TOL * ora.dt, (step, d.dt, ora.dt)
        assert abs(d.time - ora.time) <= TOL * ora.time
        assert d.step == ora.sim.step
        for bi, (bg, bo) in enumerate(zip(gpu.blocks, ora.blocks)):
            G = bo.consts["ghost"]
            errs = {
                "f": rel_err(bg.f_D.download(), bo["f"]),
                "macro": rel_err(bg.macro_D.download(), bo["macro"]),
                "Q": rel_err(bg.Q_D.download(), bo["Q"], SCALE_FLOOR["Q"]),
                "fmS": rel_err(interior_faces(bg.flux_macro_S_D.download(), G, bg.ni, bg.nj),
                               interior_faces(bo["flux_macro_S"], G, bg.ni, bg.nj)),
                "fmW": rel_err(interior_faces(bg.flux_macro_W_D.download(), G, bg.ni, bg.nj),
                               interior_faces(bo["flux_macro_W"], G, bg.ni, bg.nj)),
                "cover": rel_err(bg.wall_cover_D.download(), bo["wall_cover"]),
                "wall_fluxes": rel_err(bg.wall_fluxes_D.download(), bo["wall_fluxes"]),
                "residual": rel_err(bg.residual_D.download(), bo["residual"]),
            }
            bad = {k: v for k, v in errs.items() if not v <= TOL}
            assert not bad, (name, step, bi, bad)
    assert plan.flag() == 0.0


@pytest.mark.parametrize("name", ["cavity_vanleer_gh", "cavity_weno_nc_2x2", "cavity_weno_rho_perturbed",
                                  "cavity_weno_rho_kn0.075", "cavity_weno_rho_kn1", "cavity_weno_nv1681_rho",
                                  "cavity_weno_nv2401", "channel_adsorbing", "channel_adsorbing_cl",
                                  "channel_adsorbing_table", "channel_adsorbing_specular_sticking",
                                  "north_outflow_south_inflow", "join_self_flips", "join_NS_flipped",
                                  "curvilinear_weno_2x2", "join_N_W", "join_E_E", "cavity_weno_nv1089",
                                  "thin_two_rows", "bounce_back"])
def test_fused_mixed_error_vs_oracle(hip, name):
    gpu, plan, ora = build(name, "mixed")
    nsteps = 20
    plan.run(nsteps)
    ora.run(nsteps)
    d = plan.state()
    assert d.step == ora.sim.step
    assert abs(d.dt - ora.dt) <= 1e-5 * ora.dt
    worst = field_errors([bg.macro_D.download() for bg in gpu.blocks], [bo["macro"] for bo in ora.blocks])
    print("mixed errors", worst)
    for k in ("rho_linf", "U_linf", "T_linf"):
        assert worst[k] <= 1e-5, (name, k, worst[k])
    assert plan.flag() == 0.0


def test_fused_split_2x2_bitwise(hip):
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    one = UGKSim(cases.cavity(16, ("NC", 8), "WENO5", 0.01, n_cells_j=12))
    four = UGKSim(cases.cavity(16, ("NC", 8), "WENO5", 0.01, n_cells_j=12, nb=(2, 2)))
    p1, p4 = StepPlan(one, "fp64"), StepPlan(four, "fp64")
    p1.run(3)
    p4.run(3)
    assert p1.state().dt == p4.state().dt
    m1 = one.blocks[0].macro()
    f1 = one.blocks[0].f()
    G = one.gdata.ghost
    for b, (i0, j0) in zip(four.blocks, [(0, 0), (0, 6), (8, 0), (8, 6)]):
        assert np.array_equal(b.macro(), m1[i0:i0 + 8, j0:j0 + 6])
        assert np.array_equal(b.f()[G:-G, G:-G], f1[G + i0:G + i0 + 8, G + j0:G + j0 + 6])


def test_fused_batch_of_cases_in_one_plan(hip):
    """Independent cases of uneven size stepped by ONE plan (one launch per
    kernel per step for all of them) each keep their own time step and agree
    with the oracle stepping each case on its own."""
    from oracle import OracleSim
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    gs = cases.cfg4_batch(ncases=3, sides=(8, 12), vgrids=(8, 12))
    sims = [UGKSim(copy.deepcopy(g)) for g in gs]
    plan = StepPlan(sims, "fp64")
    oras = [OracleSim(copy.deepcopy(g), threads=4) for g in gs]
    for _ in range(3):
        plan.run(1)
        for o in oras:
            o.one_step(True)
        states = plan.states()
        for d, o in zip(states, oras):
            assert abs(d.dt - o.dt) <= TOL * o.dt
            assert abs(d.time - o.time) <= TOL * o.time
    assert len({round(d.dt, 15) for d in states}) == 3      # the cases really differ
    for s, o in zip(sims, oras):
        assert rel_err(s.blocks[0].macro(), o.blocks[0]["macro"]) <= TOL
        assert rel_err(s.blocks[0].f(), o.blocks[0]["f"]) <= TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replays_match_eager_steps(hip, precision):
    """The step captured as a HIP graph and replayed gives bit-for-bit the
    state of the eagerly launched step, replay after replay, with eager work
    and synchronisations in between."""
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    g = cases.cavity(16, ("NC", 8), "WENO5", 0.01, nb=(2, 1))
    eager_sim, graph_sim = UGKSim(copy.deepcopy(g)), UGKSim(copy.deepcopy(g))
    eager = StepPlan(eager_sim, precision)
    graph = StepPlan(graph_sim, precision, graph=True)
    for rep in range(4):
        eager.run(1 + rep % 2)
        graph.run(1 + rep % 2)
        de, dg = eager.state(), graph.state()
        assert (de.dt, de.time, de.step) == (dg.dt, dg.time, dg.step), rep
        for be, bg in zip(eager_sim.blocks, graph_sim.blocks):
            assert np.array_equal(be.macro(), bg.macro()), rep
            assert np.array_equal(eager.f64(be), graph.f64(bg)), rep
            assert np.array_equal(be.Q_D.download(), bg.Q_D.download()), rep


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotune_keeps_state_and_results(hip, precision):
    """Autotuning the face passes leaves the state alone, and the tuned plan
    steps bit for bit like the untuned one (the workgroup size only groups
    independent faces)."""
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    g = cases.cavity(16, ("NC", 8), "WENO5", 0.01, nb=(2, 1))
    s1, s2 = UGKSim(copy.deepcopy(g)), UGKSim(copy.deepcopy(g))
    p1, p2 = StepPlan(s1, precision), StepPlan(s2, precision, graph=True)
    p1.run(2)
    p2.run(2)
    f_before = [p2.f64(b) for b in s2.blocks]
    choice = p2.autotune(reps=2)
    assert set(choice.values()) <= {1, 2, 4}
    assert all(np.array_equal(a, p2.f64(b)) for a, b in zip(f_before, s2.blocks))
    p1.run(3)
    p2.run(3)
    assert p1.state().dt == p2.state().dt
    for b1, b2 in zip(s1.blocks, s2.blocks):
        assert np.array_equal(b1.macro(), b2.macro())
        assert np.array_equal(p1.f64(b1), p2.f64(b2))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_steps_do_not_depend_on_call_chunking(hip, precision):
    """Four steps in one call equal one step and then three: each call
    rederives the cell moments from f, and they must equal, bit for bit, the
    moments the update kept (a batch of three cavities whose wall
    temperatures differ, where a fused multiply-add in the update's moment
    sum once made them differ in the last bit)."""
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim

    def run(chunks):
        sims = [UGKSim(g) for g in cases.cfg4_batch(ncases=3, sides=(8, 12), vgrids=(8, 12))]
        plan = StepPlan(sims, precision)
        for c in chunks:
            plan.run(c)
        out = [(b.macro(), plan.f64(b), b.Q_D.download()) for s in sims for b in s.blocks]
        plan.close()
        return out

    for a, b in zip(run([4]), run([1, 3])):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_geometry_recomputed_in_place_under_a_plan(hip):
    """UGKSim.initGeom / UGKSBlock.updateGeom run cellGeom again into the
    same buffers, so a plan built before keeps stepping, bit for bit as one
    whose geometry was never touched."""
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    g = cases.cavity(12, ("NC", 8), "WENO5", 0.02, n_cells_j=10, nb=(2, 1))
    a, b = UGKSim(copy.deepcopy(g)), UGKSim(copy.deepcopy(g))
    pa, pb = StepPlan(a, "fp64"), StepPlan(b, "fp64")
    pa.run(2)
    pb.run(2)
    ptrs = [(x.normal_D.ptr, x.area_D.ptr) for x in b.blocks]
    area = [x.area_D.download(x.stream) for x in b.blocks]
    b.initGeom()
    b.initGeom()
    for x in b.blocks:
        x.updateGeom()
    assert [(x.normal_D.ptr, x.area_D.ptr) for x in b.blocks] == ptrs
    # (corner ghost cells have no geometry: NaN there, as in the reference)
    assert all(np.array_equal(r, x.area_D.download(x.stream), equal_nan=True) for r, x in zip(area, b.blocks))
    pa.run(2)
    pb.run(2)
    for xa, xb in zip(a.blocks, b.blocks):
        assert np.array_equal(xa.macro(), xb.macro())
        assert np.array_equal(xa.f(), xb.f())


def test_fused_mixed_batch_of_uneven_cases_vs_oracle(hip):
    """The headline form at small size: eight independent cavities of uneven
    size, Knudsen number, lid speed and wall temperature, with every
    velocity grid of the 64-case batch (17^2, 25^2, 33^2 and 41^2), stepped
    by ONE mixed-precision plan replayed as a HIP graph; after 20 steps each
    case's rho, U, T are within 1e-5 (L-inf relative) of the FP64 oracle
    stepping that case alone, and each keeps its own time step."""
    from oracle import OracleSim
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    gs = cases.cfg4_batch(ncases=8, sides=(8, 12, 16), vgrids=(16, 24, 32, 40), seed=7)
    nvs = sorted({(int(g.u_num) // 4) * 4 + 1 for g in gs})
    print(f"[gpu] mixed batch of {len(gs)} cases, velocity grids {nvs}", flush=True)
    assert nvs == [17, 25, 33, 41]
    sims = [UGKSim(copy.deepcopy(g)) for g in gs]
    plan = StepPlan(sims, "mixed", graph=True)
    oras = [OracleSim(copy.deepcopy(g), threads=4) for g in gs]
    nsteps = 20
    plan.run(nsteps)
    for o in oras:
        o.run(nsteps)
    states = plan.states()
    assert len({round(d.dt, 12) for d in states}) == len(gs)
    for i, (s, o, d) in enumerate(zip(sims, oras, states)):
        assert d.step == o.sim.step
        assert abs(d.dt - o.dt) <= 1e-5 * o.dt
        worst = field_errors([b.macro_D.download() for b in s.blocks], [b["macro"] for b in o.blocks])
        for k in ("rho_linf", "U_linf", "T_linf"):
            assert worst[k] <= 1e-5, (i, k, worst[k])
    assert plan.flag() == 0.0


def test_fused_fp64_full_size_cavity_2x2x128(hip):
    """The 2x2-block cavity at its full size (four 128^2 blocks, 33^2
    velocities, WENO5, Kn 0.01): three FP64 steps of the graph-replayed plan
    against the oracle on every host thread, f and macro within 1e-12."""
    import os

    from oracle import OracleSim
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    g = cases.cfg2_cavity_2x2()
    print("[gpu] fused fp64 full-size 2x2x128^2 cavity, 33^2 velocities", flush=True)
    gpu = UGKSim(copy.deepcopy(g))
    plan = StepPlan(gpu, "fp64", graph=True)
    ora = OracleSim(copy.deepcopy(g), threads=max(1, min(16, len(os.sched_getaffinity(0)))))
    for step in range(3):
        plan.run(1)
        ora.one_step(True)
        d = plan.state()
        assert abs(d.dt - ora.dt) <= TOL * ora.dt
    for bi, (bg, bo) in enumerate(zip(gpu.blocks, ora.blocks)):
        assert rel_err(bg.macro_D.download(), bo["macro"]) <= TOL, bi
        assert rel_err(bg.f_D.download(), bo["f"]) <= TOL, bi
        assert rel_err(bg.Q_D.download(), bo["Q"], SCALE_FLOOR["Q"]) <= TOL, bi
    assert plan.flag() == 0.0
