"""Host counterpart of ugks_sim.UGKSim: block set-up and the time loop.

Follows ugks_sim.py:52-115 (initialisation order), :274-321 (time step and
BC refresh), :407-501 (one_step, including the residual and mass monitors)
and :503-711 (the step loop: residual cadence, stop criteria, clock limit,
the runtime control file, the final-time clamp and the save cadence of
gdata.save_options), with HDF5 output and restart (ugks_sim.py:824-1117;
the writer is pyugks_amd.output).  Not rebuilt: plotting, the OpenCL
work-size search (the fused plan's autotuner takes its place), and the exec
of user scripts (the runtime control file is read as plain
``name = literal`` assignments instead).
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import bc as bcd
from .block import UGKSBlock
from .data import UGKSData, ghost_coordinates


class UGKSim:
    def __init__(self, gdata: UGKSData, stream=None, prepared: bool = False, restart: str | None = None):
        """The initialisation order of ugks_sim.py:52-115: restart, blocks
        and their geometry, initial distributions, ghost cells, first time
        step, then the initial save.  ``restart`` (or gdata.restart) names the
        global HDF5 file of a run to continue."""
        if not prepared:
            gdata.prepare()
        self.gdata = gdata
        self.stream = stream
        self.step = 0
        if restart:
            gdata.restart = restart
        self.runtime_variables()
        self.restart()
        self.force_time = False
        self.is_open = False
        self.last_error = None
        self.load_blocks()
        self.check_closed_domain()
        self.initGeom()
        self.initFunctions()
        self.updateAllBC()
        self.calcTimeStep()
        self.updateTimeStep(limit=False)
        self.saved = False
        self.output = None
        if not gdata.restart:
            self.saveToFile(save_f=gdata.save_options.save_initial_f)

    def runtime_variables(self):
        """ugks_sim.py:119-126"""
        self.time_history_residual = []
        self.time_history_residual_N = []

    def restart(self):
        """ugks_sim.py:824-838: the step, residual history and per-block state
        of the run named by gdata.restart (read by pyugks_amd.output)."""
        self.restart_data = None
        if self.gdata.restart:
            from .output import read_restart
            self.restart_data = read_restart(self.gdata.restart)
            self.step = self.gdata.step = self.restart_data["step"]
            if self.gdata.residual_options.get_residual and "residual_N" in self.restart_data:
                self.time_history_residual_N = list(self.restart_data["residual_N"])
                self.time_history_residual = list(self.restart_data["residual"])

    def load_blocks(self):
        """ugks_sim.py:239-254; each block builds its geometry as it is made
        (the ghost vertices of every block come first, as initGeom needs them)."""
        g = self.gdata
        xys = ghost_coordinates(g.blocks, g.ghost)
        self.blocks = [UGKSBlock(i, g, b, xys[i], self.stream) for i, b in enumerate(g.blocks)]
        for b in self.blocks:
            b.blockList = self.blocks

    def initGeom(self):
        """ugks_sim.py:256-262: the blocks' geometry (cellGeom) from their
        ghost-extended vertices; load_blocks has already run it, so this only
        re-runs it when called again."""
        if getattr(self, "_geom_done", False):
            xys = ghost_coordinates(self.gdata.blocks, self.gdata.ghost)
            for b in self.blocks:
                b.initGeom(xys[b.id])
        self._geom_done = True

    def initFunctions(self):
        """ugks_sim.py:264-272"""
        for b in self.blocks:
            b.initFunctions(self.restart_data["blocks"][b.id] if self.restart_data else None)

    def start_timer(self):
        """ugks_sim.py:171-178"""
        self.timer = time.time()

    def stop_timer(self):
        """ugks_sim.py:180-185"""
        return time.time() - self.timer

    @staticmethod
    def secToTime(seconds):
        """ugks_sim.py:813-822: h:mm:ss:ms"""
        s, ms = divmod(seconds * 1000, 1000)
        m, s = divmod(s, 60)
        h, m = divmod(m, 60)
        return "%d:%02d:%02d:%02d" % (h, m, s, ms)

    def updateHost(self, getF=False):
        """ugks_sim.py:713-720"""
        for b in self.blocks:
            b.updateHost(getF)

    def run_optimisation(self, plan=None, reps: int = 3):
        """ugks_sim.py:352-405 times the OpenCL work-group sizes; the fused
        step's counterpart times the face passes' workgroup sizes instead
        (StepPlan.autotune).  Without a plan there is nothing to tune: the
        reference-shaped kernels derive their launch shapes themselves."""
        return plan.autotune(reps) if plan is not None else {}

    # ---- HDF5 output (ugks_sim.py:842-1117) -------------------------------
    def h5_path(self):
        """Where initHDF puts the global file (ugks_sim.py:846-873)."""
        g, save = self.gdata, self.gdata.save_options
        if save.h5Name:
            return save.h5Name
        if g.restart:
            from .output import restart_output_name
            return restart_output_name(g.restart, g.time)
        name = save.save_name or os.path.split(os.path.abspath(g.rootName))[1]
        return os.path.join(g.rootName, "HDF5", name, name + ".h5")

    def initHDF(self):
        """ugks_sim.py:842-1020"""
        from .output import HDFOutput
        save = self.gdata.save_options
        self.output = HDFOutput(self, self.h5_path(), compression=save.compression, title=save.save_name,
                                restart_from=self.gdata.restart or None)
        return self.output

    def saveHDF5(self, save_f=False, save_flux=False, save_wall_f=False):
        """ugks_sim.py:1022-1063 (save_wall_f writes the wall distributions
        get_interface_distribution took).  While a mixed precision plan steps
        the run, f lives in FP32: it is widened into the blocks' FP64 arrays
        first (the saved heat flux and internal data are computed from it)."""
        if self.output is None:
            self.initHDF()
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.precision == "mixed":
            for b in self.blocks:
                b.f_D.upload(plan.f64(b), self.stream)
        return self.output.save(save_f=save_f, save_flux=save_flux, save_wall_f=save_wall_f)

    def close_HDF5(self):
        """ugks_sim.py:1065-1096"""
        if self.output is not None:
            self.output.close(self.time_history_residual_N, self.time_history_residual)
            self.output = None

    def saveToFile(self, close_file=False, save_f=False, save_wall_flux=False, save_wall_f=False):
        """ugks_sim.py:1098-1117: one step file per step at most (``saved``),
        only when gdata.save_options.save is set; close_file closes the run."""
        if self.gdata.save_options.save and not self.saved:
            self.saveHDF5(save_f, save_wall_flux, save_wall_f)
            self.saved = True
        if close_file:
            self.close_HDF5()

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
            self._residual_monitor()
        g.time += g.dt
        self.step += 1
        g.step = self.step
        self.updateAllBC()
        return True, False

    def _residual_monitor(self):
        """ugks_sim.py:462-488: mean block residual, plus |boundary mass flow|
        and the relative change of the total (bulk + adsorbed) mass for open
        domains."""
        g = self.gdata
        res = sum(np.asarray(b.residual, dtype=np.float64) for b in self.blocks) / len(self.blocks)
        if self.is_open:
            mdot = sum(b.bc_mdot for b in self.blocks)
            res = np.append(res, [np.nan if mdot == 0.0 else abs(mdot)])
            bulk, ads = self.total_mass()
            m0 = getattr(self, "initial_mass_total", bulk + ads)
            dm = abs((m0 - (bulk + ads)) / m0)
            res = np.append(res, [dm if dm > 0.0 else np.nan])
        g.residual_options.global_residual = res
        g.residual_options.residual_history.append(res)
        g.global_residual = res
        return res

    def analyse_residual(self):
        """ugks_sim.py:722-777 without the plot: the slope of each residual
        component on a log-log scale, from the resampled history after a
        second-order Butterworth low-pass (Wn = 1 / n^0.8) run forwards and
        backwards; inf until slope_start."""
        from scipy import signal
        res = self.gdata.residual_options
        if self.step < res.slope_start:
            return np.inf
        res_x = np.log10(np.asarray(self.time_history_residual_N, dtype=np.float64))
        res_y = np.log10(np.asarray(self.time_history_residual, dtype=np.float64))
        dx = np.min(np.diff(res_x))
        sample_x = np.arange(res_x[0], res_x[-1] + dx, dx)
        n = len(res_x)
        A, B = signal.butter(2, 1.0 / n ** 0.8, output="ba")
        slopes = []
        for i in range(res_y.shape[1]):
            ry = res_y[:, i]
            mask = np.isfinite(ry)
            filt = signal.filtfilt(A, B, np.interp(sample_x, res_x[mask], ry[mask]))
            slopes.append(np.abs(np.diff(filt) / dx)[-1])
        return slopes

    def read_runtime_conf(self):
        """The per-step re-read of gdata.runtime_conf (ugks_data.py:373-380),
        restricted to ``name = literal`` lines for the step-loop controls."""
        import ast
        g = self.gdata
        path = g.runtime_conf
        if not path:
            return
        try:
            with open(path) as fh:
                text = fh.read()
        except OSError:
            return
        allowed = {"exit", "max_time", "max_step", "print_count", "CFL", "suggest_dt", "delta_dt",
                   "dt_update_count", "clock_time_stop"}
        for line in text.splitlines():
            line = line.split("#", 1)[0].strip()
            if "=" not in line:
                continue
            key, val = (x.strip() for x in line.split("=", 1))
            if key.startswith("gdata."):
                key = key[6:]
            if key in allowed:
                setattr(g, key, ast.literal_eval(val))

    def _plan_step(self, plan, get_res):
        """One step through a fused plan (dt re-evaluated every step)."""
        g = self.gdata
        plan.run(1)
        d = plan.state()
        if d.overshoot:
            raise RuntimeError("Overshot on first iteration, decrease time step")
        if plan.flag():
            self.last_error = "error flag"
            return True
        if get_res:
            for b in self.blocks:
                b.compute_residual()
            self._residual_monitor()
        return False

    def run(self, nsteps=None, max_time=None, plan=None, log=print):
        """The step loop of ugks_sim.py:503-678.

        Stops, with ``self.exit_reason`` set, on: gdata.exit (runtime control
        file), a NaN, the time limit, the clock limit, a residual slope below
        min_slope twice in a row, the residual drop reaching res_drop, every
        residual component below min_residual, or the step count (nsteps
        more steps, else gdata.max_step).  The last step is shortened to land
        on max_time exactly (force_time).  With ``plan`` (a fused StepPlan
        over this simulation) the steps run on the device; the time step is
        then re-evaluated every step (dt_update_count = 1).
        """
        g = self.gdata
        res = g.residual_options
        max_step = self.step + nsteps if nsteps is not None else g.max_step
        t_end = g.max_time if max_time is None else max_time
        if plan is not None and g.dt_update_count != 1:
            raise ValueError("a fused plan re-evaluates dt every step (dt_update_count must be 1)")
        clock = g.clock_time_stop[0] * 3600.0 + g.clock_time_stop[1] * 60.0
        mag = 1
        below = 0
        first_res = True
        initial_res = None
        save = g.save_options
        self._plan = plan
        if g.restart:
            self.saved = True
        if save.save and not self.saved:
            self.saveToFile(save_f=save.save_initial_f)
            self.saved = False
        bulk, ads = self.total_mass()
        self.initial_mass_total = bulk + ads
        self.exit_reason = None
        t0 = time.time()
        while self.step < max_step:
            if res.get_residual and res.non_linear_output and (self.step % mag == 0) \
                    and (self.step < res.non_linear_output_limit):
                mag = int(10 ** np.floor(np.log10(self.step + 1)))
                res.residual_count = mag
                if res.non_linear_save:
                    save.save_count = mag
                if res.non_linear_dt:
                    g.dt_update_count = mag
            get_res = bool(res.get_residual and ((self.step + 1) % res.residual_count == 0))
            if g.dt_update_count < 0:
                get_dt = self.step == -g.dt_update_count
            else:
                get_dt = not (self.step + 1) % g.dt_update_count
            if plan is None:
                _, err = self.one_step(get_dt, get_res)
            else:
                err = self._plan_step(plan, get_res)
            self.saved = False
            self.read_runtime_conf()
            if g.exit:
                self.exit_reason = "interrupted"
                break
            if err:
                self.exit_reason = "NaN encountered"
                break
            if g.time >= t_end:
                self.exit_reason = "time limit reached"
                break
            if (time.time() - t0) >= clock:
                self.exit_reason = "clock time limit reached"
                break
            if get_res:
                self.time_history_residual.append(res.global_residual)
                self.time_history_residual_N.append(self.step)
                slope = self.analyse_residual()
                if self.step >= res.slope_start:
                    below = below + 1 if np.all(np.asarray(slope) < res.min_slope) else 0
                    if below >= 2:
                        self.exit_reason = "residual stabilised"
                        break
                if self.step >= res.residual_start:
                    if first_res:
                        initial_res = res.global_residual
                        first_res = False
                    elif np.all(initial_res / res.global_residual >= res.res_drop):
                        self.exit_reason = "residual reduction limit reached"
                        break
                    if np.all(res.global_residual <= res.min_residual):
                        self.exit_reason = "minimum residual reached"
                        break
            if log and g.print_count > 0 and (not self.step % g.print_count or get_res):
                log("step %d, t = %0.5f" % (self.step, g.time))
            # save cadence (ugks_sim.py:652-669)
            if save.save_count > 0 and not self.step % save.save_count and self.step <= g.max_step - 1:
                self.saveToFile(save_f=save.save_f_always)
            if g.time < save.initial_save_cutoff_time and save.initial_save_count \
                    and not self.step % save.initial_save_count:
                self.saveToFile()
            if (t_end - g.time) < g.dt:
                g.dt = t_end - g.time
                self.force_time = True
                if plan is not None:
                    plan.set_dt(g.dt)
        if self.exit_reason is None:
            self.exit_reason = "step count reached"
        self.wall_time = time.time() - t0
        # the reconstructed wall distributions, then the final save that closes the run's files
        # (ugks_sim.py:692-698); a mixed-precision plan's f is widened into the blocks' FP64 f first
        if save.save_final_wall_distribution:
            plan = getattr(self, "_plan", None)
            for b in self.blocks:
                if plan is not None and plan.precision == "mixed":
                    b.f_D.upload(plan.f64(b), self.stream)
                b.get_interface_distribution()
        self.saveToFile(close_file=True, save_f=save.save_final_f, save_wall_flux=save.save_final_flux,
                        save_wall_f=save.save_final_wall_distribution)
        return self.exit_reason in ("step count reached", "time limit reached", "residual stabilised",
                                    "residual reduction limit reached", "minimum residual reached")

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
