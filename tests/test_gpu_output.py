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
