"""Host counterpart of ugks_block.UGKSBlock, driving libugks_hip.so.

Method names, buffer names and call order follow ugks_block.py; the
pyopencl launches ``self.prg.<kernel>(queue, global, local, *args)`` become
``self.prg.<kernel>(stream, *args)`` on the HIP kernels of the same name
(launch shapes are derived inside the library).  All buffers are FP64 device
arrays in the reference layout (ugks_block.py:212-255, :338-407).
"""
from __future__ import annotations

import numpy as np

from . import bc as bcd
from . import lib as L
from .data import UGKSData, BlockDef, initial_macro
from .udf import para_table

SOUTH, WEST = np.int32(0), np.int32(1)


class UGKSBlock:
    """One block of the simulation (ugks_block.py:46-1623, step-relevant part)."""

    def __init__(self, block_id: int, gdata: UGKSData, bdef: BlockDef, xy: np.ndarray, stream=None):
        self.id = block_id
        self.gdata = gdata
        self.bdef = bdef
        self.stream = stream
        self.ghost = gdata.ghost
        self.ni, self.nj = bdef.ni, bdef.nj
        self.Nv = gdata.Nv
        self.Ni = self.ni + 2 * self.ghost
        self.Nj = self.nj + 2 * self.ghost
        self.bc_list = bdef.bc_list
        self.fill_condition = bdef.fill_condition
        self.residual = 1e10
        self.max_dt = 0.0
        self.has_accommodating = False
        self.blockList = []
        # block_CL / update_CL (ugks_block.py:151-189): constants + tables -> device
        self.consts = gdata.block_constants(bdef)
        desc, keep = L.make_desc(self.consts, gdata.velocity_tables())
        self.prg = L.Program(desc, keep)
        self.initGeom(xy)

    # ------------------------------------------------------------ buffers
    def initGeom(self, xy):
        """ugks_block.py:209-255; cellGeom runs on the device (:748-751)."""
        Ni, Nj = self.Ni, self.Nj
        self.xy_H = np.ascontiguousarray(xy, dtype=np.float64)
        if getattr(self, "xy_D", None) is not None:
            # again (updateGeom): in place, so plans holding these buffers stay valid
            self.xy_D.upload(self.xy_H, self.stream)
        else:
            self.xy_D = L.DeviceArray.from_host(self.xy_H)
            self.centre_D = L.DeviceArray((Ni, Nj, 2), fill=-1.0)
            self.side_D = L.DeviceArray((Ni, Nj, 2, 2), fill=-1.0)
            self.area_D = L.DeviceArray((Ni, Nj), fill=-1.0)
            self.normal_D = L.DeviceArray((Ni, Nj, 2, 2), fill=-1.0)
            self.length_D = L.DeviceArray((Ni, Nj, 2), fill=-1.0)
            self.para_D = L.DeviceArray((max(self.ni, self.nj), 4, 4), fill=0.0)
        self.prg.cellGeom(self.stream, self.xy_D, self.area_D, self.centre_D, self.side_D,
                          self.normal_D, self.length_D)
        self.area_H = self.area_D.download(self.stream)
        length_H = self.length_D.download(self.stream)
        G = self.ghost
        self.face_lengths = [length_H[G:-G, -G, 0], length_H[-G, G:-G, 1],
                             length_H[G:-G, G, 0], length_H[G, G:-G, 1]]
        self.total_area = float(np.sum(self.area_H[G:-G, G:-G]))
        # the parametric position of every wall cell, read by wall-value expressions (ugks_block.py:776-836)
        self.para_H = para_table(self.ni, self.nj, G, length_H, np.asarray(self.bdef.x, dtype=np.float64),
                                 np.asarray(self.bdef.y, dtype=np.float64))
        self.para_D.upload(self.para_H, self.stream)

    def initFunctions(self, restart=None):
        """ugks_block.py:258-432.  ``restart``: this block's record from
        output.read_restart -- macro and Q from the file, and f when the file
        holds it (else f starts as the equilibrium of that macro state)."""
        Ni, Nj, Nv, ni, nj = self.Ni, self.Nj, self.Nv, self.ni, self.nj
        self.macro_H, self.Q_H = initial_macro(self.bdef)
        f_restart = None
        if restart is not None:
            self.gdata.time = restart["time"]
            self.macro_H[:, :, 0] = restart["rho"]
            self.macro_H[:, :, 1:3] = restart["UV"]
            self.macro_H[:, :, 3] = 1.0 / np.asarray(restart["T"])
            self.Q_H[:] = restart["Q"]
            f_restart = restart.get("f")
        self.flag_H = np.zeros(2)
        self.flag_D = L.DeviceArray.from_host(self.flag_H)
        dist = (Ni, Nj, Nv, 2)
        self.f_D = L.DeviceArray(dist, fill=0.0)
        self.flux_f_S_D = L.DeviceArray(dist, fill=0.0)
        self.flux_f_W_D = L.DeviceArray(dist, fill=0.0)
        self.sigma_D = L.DeviceArray(dist, fill=0.0)
        self.flux_macro_S_D = L.DeviceArray((Ni, Nj, 4), fill=0.0)
        self.flux_macro_W_D = L.DeviceArray((Ni, Nj, 4), fill=0.0)
        self.prim_D = L.DeviceArray((Ni, Nj, 4), fill=0.0)
        self.aL_D = L.DeviceArray((Ni, Nj, 4), fill=0.0)
        self.aR_D = L.DeviceArray((Ni, Nj, 4), fill=0.0)
        self.aT_D = L.DeviceArray((Ni, Nj, 4), fill=0.0)
        self.faceQ_D = L.DeviceArray((Ni, Nj, 2), fill=0.0)
        self.Mxi_D = L.DeviceArray((Ni, Nj, 3), fill=0.0)
        wl = max(ni, nj)
        self.wall_prop_D = L.DeviceArray((wl, 4, 4), fill=0.0)
        cover = np.zeros((wl, 4, 4))
        for bci, bcx in enumerate(self.bc_list):
            cover[:, bci, 0] = bcx.cover_initial
        self.wall_cover_D = L.DeviceArray.from_host(cover)
        self.wall_fluxes_D = L.DeviceArray((wl, 4, 8), fill=0.0)
        self.wall_dist_D = L.DeviceArray((wl, Nv, 2), fill=0.0)
        self.macro_D = L.DeviceArray.from_host(self.macro_H)
        self.Q_D = L.DeviceArray.from_host(self.Q_H)
        self.time_step_D = L.DeviceArray((ni, nj), fill=0.0)
        self.max_freq_D = L.DeviceArray((1,), fill=0.0)
        self.residual_D = L.DeviceArray((ni, nj, 4), fill=0.0)
        self.bc_mdot_D = L.DeviceArray((wl, 4), fill=0.0)
        s = self.stream
        if f_restart is not None:
            self.f_D.upload(np.ascontiguousarray(f_restart, dtype=np.float64), s)
        else:
            self.prg.initFunctions(s, self.f_D, self.macro_D, self.Q_D)
            self.prg.calcMacro(s, self.f_D, self.macro_D)
            self.prg.calcQ(s, self.f_D, self.macro_D, self.Q_D)
        self.parametricBC(force=True)

    # ------------------------------------------------------------- BCs
    def parametricBC(self, force=False):
        """ugks_block.py:603-619: the wall values at the current time, every
        step when a UDF string of the block contains the letter t."""
        if force or self.consts["time_dependent_BC"]:
            self.prg.paraBC(self.stream, self.para_D, self.wall_prop_D, np.float64(self.gdata.time))

    def ghostExchange(self, this_face, other_block, other_face, ori=1, flip=bcd.NO_FLIP):
        """ugks_block.py:435-473"""
        o = self.blockList[other_block]
        self.prg.edgeExchange(self.stream, self.f_D, self.xy_D, np.int32(this_face), o.f_D, o.xy_D,
                              np.int32(o.Ni), np.int32(o.Nj), np.int32(other_face), np.int32(ori), np.int32(flip))

    # ghost fills of one face (ugks_block.py:475-601)
    def ghostExtrapolate(self, this_face):
        """ugks_block.py:475-492"""
        self.prg.edgeExtrapolate(self.stream, self.f_D, np.int32(this_face))

    def ghostInflow(self, this_face):
        """ugks_block.py:494-516"""
        self.prg.edgeInflow(self.stream, self.f_D, np.int32(this_face), self.normal_D, self.macro_D, self.Q_D,
                            self.wall_prop_D)

    def ghostOutflow(self, this_face):
        """ugks_block.py:518-541"""
        self.prg.edgeOutflow(self.stream, self.f_D, np.int32(this_face), self.normal_D, self.macro_D, self.Q_D,
                             self.wall_prop_D)

    def ghostConstant(self, this_face):
        """ugks_block.py:543-561"""
        self.prg.edgeConstant(self.stream, self.f_D, np.int32(this_face), self.wall_prop_D)

    def ghostMirror(self, this_face, bb=False):
        """ugks_block.py:563-582: mirror (or, with bb, bounce-back) ghost cells"""
        fn = self.prg.edgeBounceBack if bb else self.prg.edgeMirror
        fn(self.stream, self.f_D, np.int32(this_face))

    def edgeConstGrad(self, this_face):
        """ugks_block.py:584-601"""
        self.prg.edgeConstGrad(self.stream, self.f_D, np.int32(this_face))

    def updateBC(self):
        """ugks_block.py:621-665"""
        for this_face, b in enumerate(self.bc_list):
            t = b.type_of_BC
            if t in (bcd.ADJACENT, bcd.PERIODIC):
                self.ghostExchange(this_face, b.other_block, b.other_face, b.orientation, b.flip_distribution)
            elif t == bcd.EXTRAPOLATE_OUT:
                self.ghostExtrapolate(this_face)
            elif t == bcd.CONSTANT:
                self.ghostConstant(this_face)
            elif t == bcd.REFLECT:
                self.ghostMirror(this_face)
            elif t == bcd.BOUNCE_BACK:
                self.ghostMirror(this_face, bb=True)
            elif t in (bcd.DIFFUSE, bcd.ADSORBING):
                self.edgeConstGrad(this_face)
                self.has_accommodating = True
            elif t == bcd.INFLOW:
                self.ghostInflow(this_face)
            elif t == bcd.OUTFLOW:
                self.ghostOutflow(this_face)

    def updateGeom(self):
        """ugks_block.py:734-751: the geometry of the block from its
        ghost-extended vertices (cellGeom on the device)."""
        self.initGeom(self.xy_H)

    # ------------------------------------------------------------- flags
    def get_flag(self):
        """ugks_block.py:840-876 (blocking read of the device flag)."""
        self.flag_H = self.flag_D.download(self.stream)
        code = self.flag_H[0]
        if code in (1.0, 1.1, 1.2, 1.3, 4.0):
            self.last_error = {1.0: "NaN encountered", 4.0: "pressure / temperature outside adsorption isotherm data"}.get(code, "NaN")
            return True
        if code in (2.0, 3.0):
            self.max_dt = min(self.max_dt, self.flag_H[1])
        return False

    def reset_flag(self):
        """ugks_block.py:878-887"""
        self.flag_D.zero(self.stream)

    # ------------------------------------------------------------- UGKS
    def wall_flux(self, bcx, wall, flux, flux_macro, dt):
        """ugks_block.py:930-983"""
        s = self.stream
        w = np.int32(wall)
        if bcx.type_of_BC == bcd.DIFFUSE:
            self.prg.accommodatingWallDist(s, self.normal_D, w, self.wall_prop_D, flux, dt)
        else:
            self.reset_flag()
            self.prg.adsorbingWall_P1(s, self.normal_D, w, self.wall_prop_D, self.wall_cover_D, self.wall_dist_D,
                                      flux, self.macro_D, dt, self.flag_D)
            if self.get_flag():
                return True
            if bcx.reflect_type == "CL":
                self.prg.adsorbingWallCL_P2(s, self.normal_D, w, self.wall_prop_D, self.wall_cover_D,
                                            self.wall_dist_D, flux, self.macro_D, dt)
            elif bcx.reflect_type == "D":
                self.prg.adsorbingWallD_P2(s, self.normal_D, w, self.wall_prop_D, self.wall_cover_D,
                                           self.wall_dist_D, flux, self.macro_D, dt)
            else:
                self.prg.adsorbingWallS_P2(s, self.normal_D, w, self.wall_prop_D, self.wall_cover_D,
                                           self.wall_dist_D, flux, self.macro_D, dt)
        self.prg.wallFlux(s, self.normal_D, self.length_D, w, flux, flux_macro, dt)
        self.prg.wallMassEnergyFluxes(s, self.normal_D, self.length_D, self.area_D, w, flux, self.wall_fluxes_D, dt)
        return False

    def _flux_pass(self, face, dt):
        """One direction of UGKS_flux (ugks_block.py:1062-1156 / :1167-1257)."""
        s, p = self.stream, self.prg
        ff = self.flux_f_S_D if face == SOUTH else self.flux_f_W_D
        fm = self.flux_macro_S_D if face == SOUTH else self.flux_macro_W_D
        top, bot = (bcd.NORTH, bcd.SOUTH) if face == SOUTH else (bcd.EAST, bcd.WEST)
        p.iFace(s, self.f_D, ff, self.sigma_D, self.centre_D, self.side_D, self.normal_D, face)
        ot = ob = 0
        if self.bc_list[top].type_of_BC in (bcd.DIFFUSE, bcd.ADSORBING):
            ot = 1
            self.wall_flux(self.bc_list[top], top, ff, fm, dt)
        if self.bc_list[bot].type_of_BC in (bcd.DIFFUSE, bcd.ADSORBING):
            ob = 1
            self.wall_flux(self.bc_list[bot], bot, ff, fm, dt)
        ob, ot = np.int32(ob), np.int32(ot)
        p.getFaceCons(s, ff, self.normal_D, face, self.prim_D, ob, ot)
        p.getAL(s, self.f_D, self.normal_D, face, self.aL_D, ob, ot)
        p.getAR(s, self.f_D, self.normal_D, face, self.aR_D, ob, ot)
        p.getPLR(s, self.centre_D, self.side_D, face, self.prim_D, self.aL_D, self.aR_D, ob, ot)
        p.initMacroFlux(s, fm, face, dt, self.prim_D, self.aL_D, self.aR_D, self.aT_D, self.Mxi_D, ob, ot)
        p.calcFaceQ(s, ff, self.prim_D, self.normal_D, self.faceQ_D, face, ob, ot)
        p.macroFlux(s, ff, self.sigma_D, fm, self.normal_D, self.length_D, face, dt, self.prim_D, self.faceQ_D, ob, ot)
        p.distFlux(s, ff, self.sigma_D, self.normal_D, self.length_D, face, dt, self.prim_D, self.aL_D,
                   self.aR_D, self.aT_D, self.Mxi_D, self.faceQ_D, ob, ot)

    def UGKS_flux(self):
        """ugks_block.py:1049-1260"""
        dt = np.float64(self.gdata.dt)
        self._flux_pass(SOUTH, dt)
        self._flux_pass(WEST, dt)
        return False

    def UGKS_update(self, get_residual=False):
        """ugks_block.py:1262-1329"""
        s, p = self.stream, self.prg
        p.calcQ(s, self.f_D, self.macro_D, self.Q_D)
        p.updateMacro(s, self.flux_macro_S_D, self.flux_macro_W_D, self.area_D, self.macro_D, self.residual_D)
        self.reset_flag()
        dt = np.float64(self.gdata.dt)
        p.UGKS_update(s, self.f_D, self.flux_f_S_D, self.flux_f_W_D, self.area_D, self.macro_D, self.Q_D,
                      self.residual_D, dt, self.flag_D)
        if self.get_flag():
            return True
        if get_residual:
            self.compute_residual()
        return False

    def compute_residual(self):
        """The residual part of ugks_block.py:1303-1327: boundary mass flow,
        then the normalised RMS change of (rho, U, V, lambda), from residual_D
        holding the previous primary variables (updateMacro stores them)."""
        s, p = self.stream, self.prg
        self.get_BC_mdot()
        p.getResidual(s, self.macro_D, self.residual_D)
        res = self.residual_D.download(s)
        macro = self.macro_D.download(s)
        sum_res = np.sum(np.sum(res, axis=0), axis=0)
        sum_avg = np.sum(np.sum(np.abs(macro), axis=0), axis=0)
        self.residual = np.sqrt(self.ni * self.nj * sum_res) / (sum_avg + np.finfo(float).tiny)
        self.residual[self.residual <= 0.0] = 1.0
        return self.residual

    def get_BC_mdot(self):
        """ugks_block.py:889-924"""
        got = False
        for this_face, b in enumerate(self.bc_list):
            if b.type_of_BC in (bcd.EXTRAPOLATE_OUT, bcd.INFLOW, bcd.OUTFLOW, bcd.CONSTANT):
                self.prg.edgeMassFlow(self.stream, self.macro_D, self.normal_D, self.length_D, self.bc_mdot_D,
                                      np.int32(this_face))
                got = True
        self.bc_mdot = float(np.sum(self.bc_mdot_D.download(self.stream))) if got else 0.0

    def getDT(self):
        """ugks_block.py:1415-1433"""
        g = self.gdata
        self.prg.clFindDT(self.stream, self.xy_D, self.area_D, self.macro_D, self.time_step_D, np.float64(g.Kn_eff))
        self.prg.time_step_max(self.stream, self.time_step_D, self.max_freq_D)
        max_freq = float(self.max_freq_D.download(self.stream)[0])
        self.max_dt = g.CFL / max_freq
        if g.suggest_dt:
            self.max_dt = min(g.suggest_dt, self.max_dt)

    def free_device(self):
        """Release every device buffer of this block now (instead of at garbage collection)."""
        for k, v in list(vars(self).items()):
            if isinstance(v, L.DeviceArray):
                v.free()

    # ------------------------------------------------------------- host data
    def macro(self):
        return self.macro_D.download(self.stream)

    def f(self):
        return self.f_D.download(self.stream)

    def Q(self, external=False):
        """calcQ on the current state (ugks_block.py:1377-1386)."""
        self.prg.calcQ(self.stream, self.f_D, self.macro_D, self.Q_D)
        q = self.Q_D.download(self.stream)
        if external:
            q = q * (4 + self.gdata.K) / 5.0
        return q

    def updateHostMacro(self):
        """ugks_block.py:1350-1363: host copies of the macro and wall state"""
        s = self.stream
        self.macro_H = self.macro_D.download(s)
        self.wall_cover_H = self.wall_cover_D.download(s)
        self.wall_fluxes_H = self.wall_fluxes_D.download(s)
        self.wall_prop_H = self.wall_prop_D.download(s)

    def getInternal(self):
        """ugks_block.py:1334-1348: Txyz_H (Tx, Ty, Tz, T) and stress_H
        (tau_xx, tau_xy, tau_yy, 0) per cell, when save_options.internal_data."""
        if not self.gdata.save_options.internal_data:
            return None
        s = self.stream
        if getattr(self, "Txyz_D", None) is None:
            self.Txyz_D = L.DeviceArray((self.ni, self.nj, 4), fill=0.0)
            self.stress_D = L.DeviceArray((self.ni, self.nj, 4), fill=0.0)
        self.prg.getInternal(s, self.f_D, self.Txyz_D, self.stress_D)
        self.Txyz_H = self.Txyz_D.download(s)
        self.stress_H = self.stress_D.download(s)
        return self.Txyz_H, self.stress_H

    def get_interface_distribution(self):
        """ugks_block.py:985-1041: the reconstructed distribution at every wall
        face (iFace into the face buffers, whose contents are put back after):
        wall_f_N/E/S/W, [wall length][Nv][2]."""
        if not self.gdata.save_options.save_final_wall_distribution:
            return
        s, G = self.stream, self.ghost
        keep_S, keep_W = self.flux_f_S_D.download(s), self.flux_f_W_D.download(s)
        self.prg.iFace(s, self.f_D, self.flux_f_S_D, self.sigma_D, self.centre_D, self.side_D, self.normal_D,
                       np.int32(SOUTH))
        fS = self.flux_f_S_D.download(s)
        self.wall_f_N = np.copy(fS[G:self.Ni - G, self.Nj - G])
        self.wall_f_S = np.copy(fS[G:self.Ni - G, G])
        self.prg.iFace(s, self.f_D, self.flux_f_W_D, self.sigma_D, self.centre_D, self.side_D, self.normal_D,
                       np.int32(WEST))
        fW = self.flux_f_W_D.download(s)
        self.wall_f_E = np.copy(fW[self.Ni - G, G:self.Nj - G])
        self.wall_f_W = np.copy(fW[G, G:self.Nj - G])
        self.flux_f_S_D.upload(keep_S, s)
        self.flux_f_W_D.upload(keep_W, s)

    def updateHost(self, get_f=False, get_wall_f=False):
        """ugks_block.py:1350-1412: host copies of the macro and wall state
        (macro_H, wall_cover_H, wall_fluxes_H, wall_prop_H) and of the heat
        flux in external scaling (Q_H, calcQ then x(4+K)/5); returns f and/or
        the two face-flux arrays on request, as the reference does."""
        s = self.stream
        self.updateHostMacro()
        self.Q_H = self.Q(external=True)
        out = []
        if get_f:
            out.append(self.f_D.download(s))
        if get_wall_f:
            out += [self.flux_f_S_D.download(s), self.flux_f_W_D.download(s)]
        if not out:
            return None
        return out[0] if len(out) == 1 else tuple(out)

    def block_mass(self):
        """ugks_block.py:1598-1623"""
        G = self.ghost
        macro = self.macro()
        bulk = float(np.sum(self.area_H[G:-G, G:-G] * macro[:, :, 0])) * self.gdata.D_ref * self.gdata.L_ref ** 2
        cover = self.wall_cover_D.download(self.stream)
        ads = 0.0
        lens = [self.ni, self.nj, self.ni, self.nj]
        for face, b in enumerate(self.bc_list):
            ads += float(np.sum(cover[0:lens[face], face, 0] * b.S_T * self.face_lengths[face] * self.gdata.L_ref))
        return bulk, ads
