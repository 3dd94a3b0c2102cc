"""Host counterpart of ugks_sim.UGKSim: block set-up and the time loop.

Follows ugks_sim.py:52-115 (initialisation order), :274-321 (time step and
BC refresh), :407-501 (one_step) and :503-680 (the step loop, without file
output, plotting, the residual-slope analysis or the runtime .conf re-read).
"""
from __future__ import annotations

import time

import numpy as np

from . import bc as bcd
from .block import UGKSBlock
from .data import UGKSData, ghost_coordinates


class UGKSim:
    def __init__(self, gdata: UGKSData, stream=None, prepared: bool = False):
        if not prepared:
            gdata.prepare()
        self.gdata = gdata
        self.stream = stream
        self.step = 0
        self.force_time = False
        self.is_open = False
        self.last_error = None
        xys = ghost_coordinates(gdata.blocks, gdata.ghost)
        # load_blocks + initGeom (ugks_sim.py:239-262)
        self.blocks = [UGKSBlock(i, gdata, b, xys[i], stream) for i, b in enumerate(gdata.blocks)]
        for b in self.blocks:
            b.blockList = self.blocks
        self.check_closed_domain()
        # initFunctions, updateAllBC, calcTimeStep, updateTimeStep(limit=False) (ugks_sim.py:84-92)
        for b in self.blocks:
            b.initFunctions()
        self.updateAllBC()
        self.calcTimeStep()
        self.updateTimeStep(limit=False)

    def check_closed_domain(self):
        """ugks_sim.py:323-334"""
        for b in self.blocks:
            for x in b.bc_list:
                if x.type_of_BC in (bcd.EXTRAPOLATE_OUT, bcd.INFLOW, bcd.OUTFLOW, bcd.CONSTANT):
                    self.is_open = True

    def calcTimeStep(self):
        """ugks_sim.py:274-283"""
        for b in self.blocks:
            b.getDT()

    def updateTimeStep(self, limit=True):
        """ugks_sim.py:285-311"""
        g = self.gdata
        if not self.force_time:
            dt_old = g.dt
            dt_new = min(b.max_dt for b in self.blocks)
            if dt_old != dt_new:
                if (dt_new > (1.0 + g.delta_dt) * dt_old) & limit:
                    dt_new = (1.0 + g.delta_dt) * dt_old
                g.dt = dt_new

    def updateAllBC(self):
        """ugks_sim.py:313-321"""
        for b in self.blocks:
            b.updateBC()

    def one_step(self, get_dt, get_res=False):
        """ugks_sim.py:407-501; returns (step finished, error flag)."""
        g = self.gdata
        for b in self.blocks:
            b.parametricBC()
        dt_old = 0.0
        count = 0
        while (g.dt != dt_old) & (count < 2):
            dt_old = g.dt
            for b in self.blocks:
                if b.UGKS_flux():
                    return True, True
            self.updateTimeStep()
            count += 1
        if count > 1:
            if g.step == 1:
                raise RuntimeError("Overshot on first iteration, decrease time step")
            get_dt = True
        for b in self.blocks:
            if b.UGKS_update(get_res):
                self.last_error = getattr(b, "last_error", "error flag")
                return True, True
        if get_dt:
            self.calcTimeStep()
            self.updateTimeStep()
        if get_res:
            res = sum(b.residual for b in self.blocks) / len(self.blocks)
            g.global_residual = res
        g.time += g.dt
        self.step += 1
        g.step = self.step
        self.updateAllBC()
        return True, False

    def run(self, nsteps=None, max_time=None):
        """The step loop of ugks_sim.py:542-678 (stop on NaN, step or time limit)."""
        g = self.gdata
        nsteps = g.max_step if nsteps is None else nsteps
        t_end = g.max_time if max_time is None else max_time
        t0 = time.time()
        target = self.step + nsteps
        while self.step < target:
            if g.dt_update_count < 0:
                get_dt = self.step == -g.dt_update_count
            else:
                get_dt = not (self.step + 1) % g.dt_update_count
            _, err = self.one_step(get_dt)
            if err:
                return False
            if g.time >= t_end:
                break
            if (t_end - g.time) < g.dt:
                g.dt = t_end - g.time
                self.force_time = True
        self.wall_time = time.time() - t0
        return True

    def total_mass(self):
        """ugks_sim.py:1119-1131"""
        bulk = ads = 0.0
        for b in self.blocks:
            x, y = b.block_mass()
            bulk += x
            ads += y
        return bulk, ads

    def macro(self):
        return [b.macro() for b in self.blocks]
