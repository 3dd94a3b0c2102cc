"""Saving a GPU run in the reference's HDF5 layout and restarting from it
(pyugks_amd.output, UGKSim(restart=...)): the step file holds the device
state exactly, and a run restarted from it continues like the run that
wrote it.  Not bit for bit: the file keeps T, and the restart takes
lambda = 1 / T back (ugks_block.py:290), which can differ from the running
lambda in the last bit."""
import copy

import numpy as np
import pytest

from pyugks_amd import cases, h5

try:
    h5.lib()
    HAVE_H5 = True
except h5.H5Error:
    HAVE_H5 = False
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not HAVE_H5, reason="HDF5 C library not found")]


def test_save_and_restart_continue_the_run(hip, tmp_path):
    from pyugks_amd.output import HDFOutput
    from pyugks_amd.sim import UGKSim
    print("[gpu] output / restart", flush=True)
    g = cases.cavity(10, ("NC", 8), "WENO5", 0.05, lid=0.2, n_cells_j=8, nb=(2, 1))
    g.max_step = 1000
    fresh = copy.deepcopy(g)
    sim = UGKSim(g)
    for _ in range(3):
        sim.one_step(True)
    out = HDFOutput(sim, tmp_path / "HDF5" / "cav.h5")
    path = out.save(save_f=True)
    out.close()
    with h5.File(path) as d:
        for b in sim.blocks:
            blk = d["block_%d" % b.id]
            assert np.array_equal(blk["f"], b.f_D.download(b.stream))
            assert np.array_equal(blk["rho"], b.macro_D.download(b.stream)[..., 0])
            blk.close()
    rs = UGKSim(fresh, restart=str(tmp_path / "HDF5" / "cav.h5"))
    assert rs.step == sim.step and rs.gdata.time == sim.gdata.time and rs.gdata.dt == sim.gdata.dt
    for _ in range(3):
        sim.one_step(True)
        rs.one_step(True)
    for a, b in zip(rs.blocks, sim.blocks):
        fa, fb = a.f_D.download(a.stream), b.f_D.download(b.stream)
        ma, mb = a.macro_D.download(a.stream), b.macro_D.download(b.stream)
        assert np.max(np.abs(fa - fb)) <= 1e-13 * np.max(np.abs(fb))
        assert np.max(np.abs(ma - mb) / np.abs(mb).max(axis=(0, 1))) <= 1e-13


@pytest.mark.parametrize("precision", [None, "mixed"])
def test_run_saves_on_the_reference_cadence(hip, tmp_path, precision):
    """gdata.save_options drives the step loop's output as in ugks_sim.py:
    the initial state at construction, every save_count steps up to
    max_step - 1, and the final state (with f) when the loop ends, which also
    closes the run.  The files hold the device state, including f and Q (in
    external scaling) of a mixed-precision plan, and a run restarts from
    them at the final step."""
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    print(f"[gpu] save cadence, plan={precision}", flush=True)
    g = cases.cavity(8, ("NC", 8), "vanLeer", 0.05, lid=0.2, n_cells_j=6)
    g.max_step = 5
    g.rootName = str(tmp_path / "run")
    g.save_options.save = True
    g.save_options.save_count = 2
    g.save_options.save_final_f = True
    g.save_options.compression = None
    fresh = copy.deepcopy(g)
    sim = UGKSim(g)
    plan = StepPlan(sim, precision) if precision else None
    assert sim.run(plan=plan, log=None)
    assert sim.step == 5 and sim.output is None
    glob = tmp_path / "run" / "HDF5" / "run" / "run.h5"
    with h5.File(glob) as f:
        assert f["final_step"] == 5
        assert list(f["sub_files"][:, 0]) == ["run_step=%03d.h5" % k for k in (0, 2, 4, 5)]
    with h5.File(tmp_path / "run" / "HDF5" / "run" / "run_step=005.h5") as d:
        blk = d["block_0"]
        b = sim.blocks[0]
        f64 = plan.f64(b) if plan is not None else b.f_D.download(b.stream)
        assert np.array_equal(blk["f"], f64)
        assert np.array_equal(blk["Q"], b.Q(external=True))
        blk.close()
    if plan is not None:
        plan.close()
    rs = UGKSim(fresh, restart=str(glob))
    assert rs.step == 5 and rs.gdata.time == sim.gdata.time


def test_internal_data_and_wall_distributions_match_the_oracle(hip):
    """getInternal (clKernel.c:315-372) on the GPU state after three steps, and
    get_interface_distribution (ugks_block.py:985-1041): iFace into the face
    buffers, the wall rows kept, the step's fluxes put back."""
    import ctypes as C

    import oracle as O
    from helpers import rel_err
    from pyugks_amd.sim import UGKSim
    print("[gpu] getInternal / wall distributions", flush=True)
    g = cases.cavity(10, ("NC", 8), "WENO5", 0.05, lid=0.2, n_cells_j=8, nb=(2, 1), rho_perturb=0.01)
    g.save_options.internal_data = True
    g.save_options.save_final_wall_distribution = True
    sim, ora = UGKSim(copy.deepcopy(g)), O.OracleSim(copy.deepcopy(g), threads=2)
    for _ in range(3):
        sim.one_step(True)
        ora.one_step(True)
    lib = O.lib()
    lib.or_iFace.argtypes = [C.c_void_p] * 7 + [C.c_int]
    ptr = lambda a: C.c_void_p(a.ctypes.data)
    for b, bo in zip(sim.blocks, ora.blocks):
        txyz, stress = b.getInternal()
        otxyz, ostress = bo.internal()
        assert rel_err(txyz, otxyz) <= 1e-12 and rel_err(stress, ostress) <= 1e-12
        keep_S, keep_W = b.flux_f_S_D.download(), b.flux_f_W_D.download()
        b.get_interface_distribution()
        assert np.array_equal(b.flux_f_S_D.download(), keep_S) and np.array_equal(b.flux_f_W_D.download(), keep_W)
        G, Ni, Nj = b.ghost, b.Ni, b.Nj
        faces = []
        for face in (0, 1):
            ff, sg = np.zeros_like(bo["f"]), np.zeros_like(bo["f"])
            lib.or_iFace(C.byref(bo.block.p), ptr(bo["f"]), ptr(ff), ptr(sg), ptr(bo["centre"]), ptr(bo["mid_side"]),
                         ptr(bo["normal"]), face)
            faces.append(ff)
        want = {"N": faces[0][G:Ni - G, Nj - G], "S": faces[0][G:Ni - G, G], "E": faces[1][Ni - G, G:Nj - G],
                "W": faces[1][G, G:Nj - G]}
        for k, w in want.items():
            got = getattr(b, "wall_f_" + k)
            assert got.shape == (b.ni if k in "NS" else b.nj, b.Nv, 2)
            assert rel_err(got, w) <= 1e-12, k


@pytest.mark.parametrize("precision", [None, "mixed"])
def test_final_save_writes_internal_data_and_wall_distributions(hip, tmp_path, precision):
    """With internal_data and save_final_wall_distribution set, each step file
    holds Txyz (Tx, Ty, Tz) and stress (tau_xx, tau_xy, tau_yy) per cell, and
    the final one the reconstructed wall distribution of every face, as the
    blocks compute them from the saved f."""
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    print(f"[gpu] internal data / wall_f in the step files, plan={precision}", flush=True)
    g = cases.cavity(8, ("NC", 8), "vanLeer", 0.05, lid=0.2, n_cells_j=6)
    g.max_step = 4
    g.rootName = str(tmp_path / "run")
    g.save_options.save = True
    g.save_options.save_count = 2
    g.save_options.internal_data = True
    g.save_options.save_final_wall_distribution = True
    g.save_options.compression = None
    sim = UGKSim(g)
    plan = StepPlan(sim, precision) if precision else None
    assert sim.run(plan=plan, log=None)
    b = sim.blocks[0]
    txyz, stress = b.getInternal()
    with h5.File(tmp_path / "run" / "HDF5" / "run" / "run_step=004.h5") as d:
        blk = d["block_0"]
        assert np.array_equal(blk["Txyz"], txyz[:, :, 0:3]) and np.array_equal(blk["stress"], stress[:, :, 0:3])
        for k in ("N", "E", "S", "W"):
            assert np.array_equal(blk["wall"][k]["wall_f"], getattr(b, "wall_f_" + k))
        blk.close()
    with h5.File(tmp_path / "run" / "HDF5" / "run" / "run_step=002.h5") as d:
        blk = d["block_0"]
        assert blk["Txyz"].shape == (8, 6, 3) and "wall_f" not in blk["wall"]["N"]
        blk.close()
    xmf = (tmp_path / "run" / "HDF5" / "run" / "run.xmf").read_text()
    assert xmf.count('<Attribute Name="Txyz" AttributeType="Vector"') == 3
    if plan is not None:
        plan.close()
