"""The HDF5 binding and the reference-layout writer, on the CPU: typed
round trips through pyugks_amd.h5, and the global / step files and XDMF of
pyugks_amd.output written for a host-side stand-in of a simulation (the
device arrays replaced by host arrays), then read back with read_restart."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from pyugks_amd import bc as bcd
from pyugks_amd import cases, h5

try:
    h5.lib()
    HAVE_H5 = True
except h5.H5Error:
    HAVE_H5 = False
pytestmark = pytest.mark.skipif(not HAVE_H5, reason="HDF5 C library not found")


def test_h5_round_trip(tmp_path):
    p = tmp_path / "t.h5"
    a = np.random.default_rng(0).standard_normal((3, 4, 2))
    with h5.File(p, "w") as f:
        f.create_dataset("a", a, compression="gzip")
        f.create_dataset("i", np.arange(5, dtype=np.int32))
        f.create_dataset("n", 7)
        f.create_dataset("x", 0.25)
        f.create_dataset("s", "title text")
        f.create_dataset("names", np.array([["run_step=0003.h5"], ["run_step=0006.h5"]], dtype=object))
        g = f.create_group("block_0")
        w = g.require_group("wall")
        w.create_dataset("T", np.ones(4))
        w.close()
        g.close()
    with h5.File(p) as f:
        assert sorted(f.keys()) == ["a", "block_0", "i", "n", "names", "s", "x"]
        assert np.array_equal(f["a"], a) and f["i"].dtype == np.int32 and list(f["i"]) == [0, 1, 2, 3, 4]
        assert f["n"] == 7 and np.ndim(f["n"]) == 0 and f["x"] == 0.25
        assert f["s"] == "title text" and f["names"][-1, 0] == "run_step=0006.h5"
        assert "block_0/wall/T" in f and "block_0/nope" not in f
        assert np.array_equal(f["block_0"]["wall"]["T"], np.ones(4))


class _Host:
    def __init__(self, a):
        self.a = np.asarray(a)

    def download(self, stream=None):
        return self.a.copy()


class _Block:
    """The attributes HDFOutput reads from a UGKSBlock, held on the host."""

    def __init__(self, bid, gdata, bdef, rng, xy):
        self.id, self.bdef, self.stream = bid, bdef, None
        self.K = gdata.K
        self.ghost = G = gdata.ghost
        self.ni, self.nj = bdef.ni, bdef.nj
        self.Ni, self.Nj = self.ni + 2 * G, self.nj + 2 * G
        self.bc_list = bdef.bc_list
        self.consts = {"ni": self.ni, "nj": self.nj}
        self.xy_H = xy
        wl = max(self.ni, self.nj)
        self.centre_D = _Host(rng.uniform(size=(self.Ni, self.Nj, 2)))
        m = rng.uniform(0.5, 1.5, (self.ni, self.nj, 4))
        self.macro_D, self.Q_D = _Host(m), _Host(rng.standard_normal((self.ni, self.nj, 2)))
        self.wall_prop_D = _Host(rng.uniform(0.5, 1.5, (wl, 4, 4)))
        self.wall_cover_D = _Host(rng.uniform(size=(wl, 4, 4)))
        self.wall_fluxes_D = _Host(rng.standard_normal((wl, 4, 8)))
        dist = (self.Ni, self.Nj, gdata.Nv, 2)
        self.f_D = _Host(rng.uniform(size=dist))
        self.flux_f_S_D, self.flux_f_W_D = _Host(rng.standard_normal(dist)), _Host(rng.standard_normal(dist))

    def Q(self, external=False):
        return self.Q_D.download() * ((4 + self.K) / 5.0 if external else 1.0)


class _Sim:
    def __init__(self, gdata):
        gdata.prepare()
        rng = np.random.default_rng(1)
        self.gdata, self.step = gdata, 0
        from pyugks_amd.data import ghost_coordinates
        xys = ghost_coordinates(gdata.blocks, gdata.ghost)
        self.blocks = [_Block(k, gdata, b, rng, xys[k]) for k, b in enumerate(gdata.blocks)]


def test_writer_layout_xdmf_and_restart_read(tmp_path):
    from pyugks_amd.output import HDFOutput, read_restart
    sim = _Sim(cases.cavity(6, ("NC", 4), "vanLeer", 0.1, nb=(2, 1), n_cells_j=4))
    sim.gdata.max_step = 100
    out = HDFOutput(sim, tmp_path / "HDF5" / "cav.h5", compression=4, title="cavity")
    sim.step, sim.gdata.time = 3, 0.125
    first = out.save()
    sim.step, sim.gdata.time = 6, 0.25
    last = out.save(save_f=True, save_flux=True)
    out.close(residual_N=[3, 6], residual=[[1e-3, 2e-3, 3e-3, 4e-3], [1e-4, 2e-4, 3e-4, 4e-4]])
    assert os.path.basename(first) == "cav_step=003.h5" and os.path.basename(last) == "cav_step=006.h5"
    with h5.File(tmp_path / "HDF5" / "cav.h5") as g:
        assert g["Nb"] == 2 and g["final_step"] == 6 and g["title"] == "cavity"
        assert list(g["sub_files"][:, 0]) == ["cav_step=003.h5", "cav_step=006.h5"]
        assert g["residual_xy"].shape == (2, 5)
        b0 = g["block_0"]
        assert b0["x"].shape == (4, 5, 1) and b0["ghost"] == sim.gdata.ghost
        assert b0["BC"]["north"]["type_of_BC"] == "DIFFUSE" and b0["BC"]["east"]["type_of_BC"] == "ADJACENT"
    with h5.File(last) as d:
        b = d["block_1"]
        blk = sim.blocks[1]
        m = blk.macro_D.a
        assert d["step"] == 6 and b["time"] == 0.25
        assert np.array_equal(b["rho"], m[..., 0]) and np.array_equal(b["UV"], m[..., 1:3])
        # the heat flux is saved in external scaling, as updateHost leaves it
        assert np.array_equal(b["Q"], blk.Q(external=True))
        assert np.array_equal(b["T"], 1.0 / m[..., 3]) and np.array_equal(b["f"], blk.f_D.a)
        G = blk.ghost
        assert np.array_equal(b["wall"]["S"]["flux"], blk.flux_f_S_D.a[G:blk.Ni - G, G])
        assert np.array_equal(b["wall"]["E"]["mom_t_out"], blk.wall_fluxes_D.a[0:blk.nj, bcd.EAST, 5])
    # the XDMF is well formed and names every field of both step files
    root = ET.parse(tmp_path / "HDF5" / "cav.xmf").getroot()
    times = [float(t.get("Value")) for t in root.iter("Time")]
    assert times == [0.125, 0.25]
    refs = [d.text.strip() for d in root.iter("DataItem") if d.get("Format") == "HDF"]
    assert "cav_step=006.h5:/block_1/UV" in refs and "cav.h5:/block_0/x" in refs
    # and the restart reader finds the last step file and its f
    r = read_restart(str(tmp_path / "HDF5" / "cav.h5"))
    assert r["step"] == 6 and r["residual_N"] == [3.0, 6.0]
    assert np.array_equal(r["blocks"][1]["f"], sim.blocks[1].f_D.a) and r["blocks"][0]["time"] == 0.25
