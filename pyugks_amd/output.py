"""HDF5 + XDMF output and restart in the reference's file layout.

``HDFOutput`` restates UGKSim.initHDF / saveHDF5 / close_HDF5
(ugks_sim.py:842-1096) with UGKSBlock.save_hdf (ugks_block.py:1435-1595):
one global file with the run constants, the block geometry and BCs, the list
of step files and the residual history, one file per saved step with the
macroscopic fields, the wall state and optionally f and the wall fluxes, and
an XDMF time series that points into the step files.  ``read_restart``
reads such a run back the way UGKSim.restart and UGKSBlock.initFunctions do
(ugks_sim.py:824-838, ugks_block.py:275-300).

Differences from the reference, all in what is stored rather than where:
the block's "CL_source" holds a description of the HIP library (there is no
OpenCL source), and the global "device" / "platform" name the HIP device.  Needs the HDF5 C library
(pyugks_amd.h5)."""
from __future__ import annotations

import math
import os

import numpy as np

from . import bc as bcd
from . import h5

FACE_KEYS = ["N", "E", "S", "W"]
BC_NAMES = {bcd.ADJACENT: "ADJACENT", bcd.EXTRAPOLATE_OUT: "EXTRAPOLATE_OUT", bcd.CONSTANT: "CONSTANT",
            bcd.REFLECT: "REFLECT", bcd.DIFFUSE: "DIFFUSE", bcd.PERIODIC: "PERIODIC", bcd.INFLOW: "INFLOW",
            bcd.OUTFLOW: "OUTFLOW", bcd.ADSORBING: "ADSORBING", bcd.BOUNCE_BACK: "BOUNCE_BACK"}


def restart_output_name(restart_path: str, time: float) -> str:
    """The global file a restarted run writes to (ugks_sim.py:857-862)."""
    path, name = os.path.split(restart_path)
    name, ext = os.path.splitext(name)
    name = (name + "_restart_t=%g" % time).replace(".", "-")
    return os.path.join(path, name) + ext


def _cell_item(h5name, bid, field, ni, nj, dims=""):
    return ('<DataItem Dimensions="%d %d%s" NumberType="Float" Precision="8" Format="HDF">\n'
            '%s:/block_%d/%s\n</DataItem>\n' % (ni, nj, dims, h5name, bid, field))


def _vector_attribute(name, h5name, bid, field, ni, nj):
    # the vector is joined with a zero third component built from rho (ugks_block.py:1460-1491)
    return ('<Attribute Name="%s" AttributeType="Vector" Center="Cell">\n' % name +
            '<DataItem Dimensions="%d %d 3" Function="JOIN($0, $1)" ItemType="Function">\n' % (ni, nj) +
            _cell_item(h5name, bid, field, ni, nj, " 2") +
            '<DataItem Dimensions="%d %d 1" Function="ABS($0 - $0)" ItemType="Function">\n' % (ni, nj) +
            _cell_item(h5name, bid, "rho", ni, nj) + '</DataItem>\n</DataItem>\n</Attribute>\n')


class HDFOutput:
    """Writer bound to one simulation (a pyugks_amd.sim.UGKSim)."""

    def __init__(self, sim, h5_path: str, compression=None, title: str = "", restart_from: str | None = None):
        self.sim, self.compression = sim, compression
        self.h5_name = os.fspath(h5_path)
        os.makedirs(os.path.dirname(os.path.abspath(self.h5_name)), exist_ok=True)
        self.base = os.path.splitext(self.h5_name)[0]
        self.sub_names = []
        g = sim.gdata
        self.glob = h5.File(self.h5_name, "w")
        grp = self.glob
        cfg = {k: getattr(g, k) for k in ("flux_method", "CFL", "max_time", "max_step", "Kn_eff", "omega", "Pr",
                                           "relax_type", "quad_type")}
        grp.create_dataset("run_config", repr(cfg))
        grp.create_dataset("device", "HIP device (libugks_hip.so, gfx950)")
        grp.create_dataset("platform", "ROCm")
        grp.create_dataset("title", title)
        for k in ("CFL", "Nv", "Kn_eff", "omega", "R", "gamma", "Pr", "L_ref", "D_ref", "T_ref", "C_ref", "t_ref",
                  "b", "K"):
            grp.create_dataset(k, getattr(g, k))
        grp.create_dataset("t0", g.time)
        grp.create_dataset("Nb", len(sim.blocks))
        grp.create_dataset("quad", np.asarray(g.quad, dtype=np.float64))
        grp.create_dataset("weight", np.asarray(g.weight, dtype=np.float64))
        if restart_from:
            grp.create_dataset("restart", restart_from)
        short = os.path.split(self.h5_name)[1]
        self.xdmf_blocks = []
        for b in sim.blocks:
            self._block_constants(grp, b)
            ni, nj = b.ni, b.nj
            geo = ['<Grid Name="Block_%d" GridType="Uniform">\n' % b.id,
                   '<Topology TopologyType="2DSMesh" NumberOfElements="%d %d"/>\n' % (ni + 1, nj + 1),
                   '<Geometry GeometryType="X_Y">\n']
            for c in ("x", "y"):
                geo.append('<DataItem Dimensions="%d %d" NumberType="Float" Precision="8" Format="HDF">\n'
                           '%s:/block_%d/%s\n</DataItem>\n' % (ni + 1, nj + 1, short, b.id, c))
            geo.append('</Geometry>\n')
            self.xdmf_blocks.append("".join(geo))
        self.xdmf = open(self.base + ".xmf", "w")
        self.xdmf.write('<?xml version="1.0" ?>\n<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>\n<Xdmf Version="2.2">\n'
                        '<Domain>\n<Grid Name="TimeSeries" GridType="Collection" CollectionType="Temporal">\n')

    def _block_constants(self, grp, b):
        c = self.compression
        blk = grp.create_group("block_%d" % b.id)
        blk.create_dataset("CL_source", "HIP kernels of libugks_hip.so; block descriptor: " +
                           repr({k: v for k, v in b.consts.items() if np.ndim(v) == 0}))
        blk.create_dataset("label", b.bdef.label or ("block_%d" % b.id))
        for k in ("Ni", "Nj", "ni", "nj", "ghost"):
            blk.create_dataset(k, int(getattr(b, k)))
        x = np.asarray(b.bdef.x, dtype=np.float64)[..., None]
        y = np.asarray(b.bdef.y, dtype=np.float64)[..., None]
        blk.create_dataset("x", x, compression=c)
        blk.create_dataset("y", y, compression=c)
        blk.create_dataset("xy", b.xy_H, compression=c)
        G = b.ghost
        centre = b.centre_D.download(b.stream)[G:-G, G:-G]
        blk.create_dataset("centreX", centre[..., 0], compression=c)
        blk.create_dataset("centreY", centre[..., 1], compression=c)
        bcg = blk.create_group("BC")
        for face, bc in enumerate(b.bc_list):
            f = bcg.create_group(bcd.faceName[face])
            f.create_dataset("type_of_BC", BC_NAMES[bc.type_of_BC])
            for k in ("other_block", "other_face", "orientation"):
                f.create_dataset(k, int(getattr(bc, k)))
            f.create_dataset("label", getattr(bc, "label", "") or "")
            for k in ("D", "U", "V", "T", "P"):
                f.create_dataset(k, float(getattr(bc, k)))
            if bc.type_of_BC == bcd.ADSORBING:
                f.create_dataset("adsorb", np.asarray(bc.adsorb, dtype=np.float64))
                for k in ("beta_n", "beta_t"):
                    f.create_dataset(k, float(getattr(bc, k)))
                # the reference stores alpha_t under both names (ugks_sim.py:959-960)
                f.create_dataset("alpha_n", float(bc.alpha_t))
                f.create_dataset("alpha_t", float(bc.alpha_t))
                for k in ("k_f", "S_T", "alpha_p", "cover_initial"):
                    f.create_dataset(k, float(getattr(bc, k)))
                f.create_dataset("reflect_type", bc.reflect_type)
            xs, ys = x[..., 0], y[..., 0]
            edge = {0: (xs[:, -1], ys[:, -1]), 1: (xs[-1, :], ys[-1, :]), 2: (xs[:, 0], ys[:, 0]),
                    3: (xs[0, :], ys[0, :])}[face]
            f.create_dataset("x_edge", np.ascontiguousarray(edge[0]), compression=c)
            f.create_dataset("y_edge", np.ascontiguousarray(edge[1]), compression=c)
            f.close()
        bcg.close()
        blk.close()

    def save(self, save_f: bool = False, save_flux: bool = False, save_wall_f: bool = False) -> str:
        """One step file (saveHDF5, ugks_sim.py:1021-1064); returns its path."""
        sim, g, c = self.sim, self.sim.gdata, self.compression
        width = int(max(math.log10(max(g.max_step, 1)), 3))
        path = self.base + ("_step=%0." + str(width) + "i") % sim.step + ".h5"
        short = os.path.split(path)[1]
        self.sub_names.append(short)
        self.xdmf.write('<Grid Name="TimeSlice" GridType="Collection" CollectionType="Spatial">\n'
                        '<Time Value="%0.15f" />\n' % g.time)
        with h5.File(path, "w") as f:
            f.create_dataset("step", int(sim.step))
            for b in sim.blocks:
                xd = self._save_block(f, b, short, save_f, save_flux, save_wall_f)
                self.xdmf.write(self.xdmf_blocks[b.id])
                self.xdmf.write(xd)
                self.xdmf.write("</Grid>\n")
        self.xdmf.write("</Grid>\n")
        return path

    def _save_block(self, f, b, short, save_f, save_flux, save_wall_f=False):
        g, c, s = self.sim.gdata, self.compression, b.stream
        ni, nj, G = b.ni, b.nj, b.ghost
        grp = f.create_group("block_%d" % b.id)
        grp.create_dataset("dt", float(g.dt))
        grp.create_dataset("time", float(g.time))
        macro = b.macro_D.download(s)
        Q = b.Q(external=True)         # updateHost: calcQ on the state, x(4+K)/5 (ugks_block.py:1376-1386)
        grp.create_dataset("rho", macro[:, :, 0], compression=c)
        grp.create_dataset("UV", macro[:, :, 1:3], compression=c)
        grp.create_dataset("Q", Q, compression=c)
        grp.create_dataset("T", 1.0 / macro[:, :, 3], compression=c)
        xd = ['<Attribute Name="rho" AttributeType="Scalar" Center="Cell">\n' + _cell_item(short, b.id, "rho", ni, nj) +
              '</Attribute>\n', _vector_attribute("UV", short, b.id, "UV", ni, nj),
              _vector_attribute("Q", short, b.id, "Q", ni, nj),
              '<Attribute Name="T" AttributeType="Scalar" Center="Cell">\n' + _cell_item(short, b.id, "T", ni, nj) +
              '</Attribute>\n',
              '<Attribute Name="P" AttributeType="Scalar" Center="Cell">\n'
              '<DataItem Dimensions="%d %d" Function="$0*$1/2.0" ItemType="Function">\n' % (ni, nj) +
              _cell_item(short, b.id, "rho", ni, nj) + _cell_item(short, b.id, "T", ni, nj) +
              '</DataItem>\n</Attribute>\n']
        if g.save_options.internal_data:
            # getInternal (clKernel.c:315-372), written as in ugks_block.py:1493-1507
            Txyz, stress = b.getInternal()
            grp.create_dataset("Txyz", np.ascontiguousarray(Txyz[:, :, 0:-1]), compression=c)
            grp.create_dataset("stress", np.ascontiguousarray(stress[:, :, 0:3]), compression=c)
            for name in ("Txyz", "stress"):
                xd.append('<Attribute Name="%s" AttributeType="Vector" Center="Cell">\n' % name +
                          _cell_item(short, b.id, name, ni, nj, " 3") + '</Attribute>\n')
        prop = b.wall_prop_D.download(s)
        cover = b.wall_cover_D.download(s)
        wflux = b.wall_fluxes_D.download(s)
        if save_flux:
            fS, fW = b.flux_f_S_D.download(s), b.flux_f_W_D.download(s)
        wall = grp.create_group("wall")
        names = ("mass_in", "mass_out", "mom_n_in", "mom_n_out", "mom_t_in", "mom_t_out", "nrg_in", "nrg_out")
        for fid, key in enumerate(FACE_KEYS):
            n = (ni, nj, ni, nj)[fid]
            w = wall.create_group(key)
            w.create_dataset("rho", prop[0:n, fid, 0], compression=c)
            w.create_dataset("U", prop[0:n, fid, 1], compression=c)
            w.create_dataset("V", prop[0:n, fid, 2], compression=c)
            with np.errstate(divide="ignore"):
                w.create_dataset("T", 1.0 / prop[0:n, fid, 3], compression=c)
            for k, name in enumerate(("cover", "reflected", "adsorbed", "desorbed")):
                w.create_dataset(name, cover[0:n, fid, k], compression=c)
            for k, name in enumerate(names):
                w.create_dataset(name, wflux[0:n, fid, k], compression=c)
            if save_flux:
                data = {0: fS[G:b.Ni - G, b.Nj - G], 1: fW[b.Ni - G, G:b.Nj - G], 2: fS[G:b.Ni - G, G],
                        3: fW[G, G:b.Nj - G]}[fid]
                w.create_dataset("flux", np.ascontiguousarray(data), compression=c)
            if save_wall_f:
                # the reconstructed wall distribution (get_interface_distribution, ugks_block.py:1575-1588)
                w.create_dataset("wall_f", np.ascontiguousarray((b.wall_f_N, b.wall_f_E, b.wall_f_S, b.wall_f_W)[fid]),
                                 compression=c)
            w.close()
        wall.close()
        if save_f:
            grp.create_dataset("f", b.f_D.download(s), compression=c)
        grp.close()
        return "".join(xd)

    def close(self, residual_N=None, residual=None):
        """close_HDF5 (ugks_sim.py:1066-1096)."""
        if self.glob is None:
            return
        grp = self.glob
        grp.create_dataset("final_step", int(self.sim.step))
        grp.create_dataset("sub_files", np.array(self.sub_names or [""], dtype=object).reshape(-1, 1))
        if residual_N is not None and len(residual_N):
            r = np.asarray(residual, dtype=np.float64).reshape(len(residual_N), -1)
            grp.create_dataset("residual_xy", np.column_stack([np.asarray(residual_N, dtype=np.float64), r]),
                               compression=self.compression)
        self.xdmf.write("</Grid>\n</Domain>\n</Xdmf>\n")
        self.xdmf.close()
        self.glob.close()
        self.glob = None


def read_restart(path: str) -> dict:
    """The state a restart starts from (ugks_sim.py:824-838,
    ugks_block.py:275-300): the final step, the residual history, and per
    block the time, rho, U, V, T, Q and -- when the last step file holds it --
    the full f with ghosts."""
    out = {"blocks": {}}
    with h5.File(path) as g:
        out["step"] = int(g["final_step"])
        if "residual_xy" in g:
            r = g["residual_xy"]
            out["residual_N"] = r[:, 0].tolist()
            out["residual"] = r[:, 1:5].tolist()
        last = str(g["sub_files"][-1, 0])
    data_path = os.path.join(os.path.split(path)[0], last)
    with h5.File(data_path) as d:
        for key in d.keys():
            if not key.startswith("block_"):
                continue
            blk = d[key]
            rec = {"time": float(blk["time"]), "rho": blk["rho"], "UV": blk["UV"], "T": blk["T"], "Q": blk["Q"]}
            if "f" in blk:
                rec["f"] = blk["f"]
            out["blocks"][int(key.split("_")[1])] = rec
            blk.close()
    return out
