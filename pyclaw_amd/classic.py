"""Classic Clawpack solvers (pyclaw/classic/solver.py:19-706) on the GPU.

ClawSolver2D.step_hyperbolic launches the fused HIP step (unsplit, or the two
dimension-split sweeps) on the device-resident state; the ghost fill happens
inside the kernels.  ClawSolver3D takes the three dimension-split sweeps.
"""
import numpy as np

from . import kernels as K
from .limiters import tvd
from .solver import BC, Solver


class ClawSolver(Solver):
    """classic/solver.py:19-243"""

    def __init__(self, riemann_solver=None, claw_package=None):
        self.num_ghost = 2
        self.limiters = tvd.minmod
        self.order = 2
        self.source_split = 1
        self.fwave = False
        self.step_source = None
        self.kernel_language = "Fortran"
        self.verbosity = 0
        self.cfl_max = 1.0
        self.cfl_desired = 0.9
        self._mthlim = self.limiters
        self._method = None
        self.dt_old = None
        self._rp_struct = None
        super(ClawSolver, self).__init__(riemann_solver, claw_package)

    # ---- classic/solver.py:96-146
    def step(self, solution, take_one_step, tstart, tend):
        self.get_dt(solution.t, tstart, tend, take_one_step)
        self.cfl.set_global_max(0.)
        if self.source_split == 2 and self.step_source is not None:
            self.step_source(self, solution.states[0], self.dt / 2.0)
        self.step_hyperbolic(solution)
        if self.cfl.get_cached_max() >= self.cfl_max:
            return False
        if self.step_source is not None:
            if self.source_split == 2:
                self.step_source(self, solution.states[0], self.dt / 2.0)
            if self.source_split == 1:
                self.step_source(self, solution.states[0], self.dt)
        return True

    def _check_cfl_settings(self):
        pass

    def _set_mthlim(self):
        self._mthlim = self.limiters
        if not isinstance(self.limiters, list):
            self._mthlim = [self._mthlim]
        if len(self._mthlim) == 1:
            self._mthlim = self._mthlim * self.num_waves
        if len(self._mthlim) != self.num_waves:
            raise Exception("Length of solver.limiters is not equal to 1 or to solver.num_waves")

    def _set_method(self, state):
        self._method = np.empty(7, dtype=int)
        self._method[0] = self.dt_variable
        self._method[1] = self.order
        if self.num_dim == 1:
            self._method[2] = 0
        elif self.dimensional_split:
            self._method[2] = -1
        else:
            self._method[2] = self.transverse_waves
        self._method[3] = self.verbosity
        self._method[4] = 0
        self._method[5] = state.index_capa + 1
        # aux fields the Riemann solver does not read (kept for the user's
        # step_source / BC functions) are not handed to the kernels
        self._method[6] = state.num_aux if getattr(self.rp, "num_aux", 0) else 0

    def setup(self, solution):
        self._check_cfl_settings()
        self._set_mthlim()
        if self.fwave and self.num_dim != 2:
            raise NotImplementedError("f-wave Riemann solvers are supported by the 2D steps")
        state = solution.states[0]
        if self.rp is None or not hasattr(self.rp, "kind"):
            raise Exception("the solver needs one of the pyclaw_amd.riemann modules")
        rp_aux = getattr(self.rp, "num_aux", 0)
        if rp_aux and not rp_aux <= state.num_aux <= 3:
            raise NotImplementedError("the Riemann solver reads %d aux fields; the device steps "
                                      "take %d to 3, the state has %d"
                                      % (rp_aux, rp_aux, state.num_aux))
        if rp_aux and self.num_dim != 2:
            raise NotImplementedError("aux fields are supported by the 2D steps")
        if state.index_capa >= 0:
            if self.num_dim != 2 or self.dimensional_split or not rp_aux:
                raise NotImplementedError("capacity functions are supported by the unsplit 2D "
                                          "step with the variable-coefficient solvers")
            if state.index_capa >= state.num_aux:
                raise ValueError("index_capa %d is not an aux field" % state.index_capa)
        if (getattr(self.rp, "kind", 0) in (6, 9, 11) and self.num_dim == 2
                and not self.dimensional_split and self.transverse_waves != 0):
            raise NotImplementedError("%s has no transverse solver here: use dimensional "
                                      "splitting or transverse_waves = 0"
                                      % self.rp.__name__.split(".")[-1])
        self._set_method(state)
        self._rp_struct = K.make_rp(self.rp.kind, state.problem_data, meqn=state.num_eqn,
                                    fwave=self.fwave)
        # the host qbc that custom BC functions fill: kept from step to step as
        # the reference keeps solver.qbc (solver.py:341-343)
        self.qbc = None
        self._is_set_up = True

    def _custom_bc_input(self, state, lo, hi):
        """User BC functions act on a host qbc (solver.py:393-398); the kernel then
        reads the stored ghosts (BC_GIVEN)."""
        import torch
        from . import _lib as L
        from .device import to_planar, torch_dtype
        ng = self.num_ghost
        q = state.q
        shape = (q.shape[0],) + tuple(n + 2 * ng for n in q.shape[1:])
        if self.qbc is None or self.qbc.shape != shape:
            self.qbc = np.zeros(shape, order="F")
        qbc = self.qbc
        qbc[(slice(None),) + (slice(ng, -ng),) * self.num_dim] = q
        if state.num_aux > 0:
            # the user's functions get auxbc too (solver.py:371-398)
            self._host_auxbc(state)
        _fill_bcs_host(self, state, qbc, lo, hi)
        t = torch.from_numpy(to_planar(qbc)).to("cuda", torch_dtype(self._dtype_code))
        comp = state.device_q(self._dtype_code)[1]
        if comp is not None:
            cp = torch.zeros_like(t)
            cp[(slice(None),) + (slice(ng, -ng),) * self.num_dim] = comp
            comp = cp
        if self.num_dim == 3:
            return (t, comp), L.geom3(q.shape[1], q.shape[2], q.shape[3], ng)
        return (t, comp), L.geom2(q.shape[1], q.shape[2], ng)


class ClawSolver2D(ClawSolver):
    """classic/solver.py:393-549"""
    no_trans = 0
    trans_inc = 1
    trans_cor = 2

    def __init__(self, riemann_solver=None, claw_package=None):
        self.dimensional_split = True
        self.transverse_waves = self.trans_inc
        self.num_dim = 2
        self.reflect_index = [1, 2]
        self.aux1 = None
        self.aux2 = None
        self.aux3 = None
        self.work = None
        self._work = None
        # Run evolve_to_time(solution, tend) as device-controlled launches (no
        # host synchronisation per step) when no host callback is involved.
        self.device_loop = True
        self.use_graph = True
        self.graph_chunk = 16
        self._ctl = None
        self._graphs = {}
        super(ClawSolver2D, self).__init__(riemann_solver, claw_package)

    def _check_cfl_settings(self):
        if (not self.dimensional_split) and (self.transverse_waves == 0):
            cfl_recommended = 0.5
        else:
            cfl_recommended = 1.0
        if self.cfl_max > cfl_recommended:
            import warnings
            warnings.warn("cfl_max is set higher than the recommended value of %s" % cfl_recommended)
            warnings.warn(str(self.cfl_desired))

    def step_hyperbolic(self, solution):
        state = solution.states[0]
        grid = state.grid
        dx, dy = grid.delta
        lo, hi = self._bc_codes()
        qin = state.device_q(self._dtype_code)
        out = self._take_scratch(qin)
        cfl_dev = self._cfl_buffer()
        K.cfl_reset(cfl_dev)
        gin = None
        src = qin
        aux = alo = ahi = None
        if state.num_aux > 0 and getattr(self.rp, "num_aux", 0):
            aux, alo, ahi = self._aux_input(state)
        custom = BC.custom in lo or BC.custom in hi
        if custom:
            src, gin = self._custom_bc_input(state, lo, hi)
            lo, hi = [4, 4], [4, 4]
        if self.dimensional_split and custom:
            # the sweeps write the whole qbc, ghost rows and columns included,
            # which the next step's BC functions see where they set nothing
            import torch
            ng = self.num_ghost
            whole = (torch.empty_like(src[0]),
                     torch.empty_like(src[1]) if src[1] is not None else None)
            K.step2_split(self._rp_struct, src[0], whole[0], cfl_dev, lo, hi,
                          self.reflect_index, self._method, self._mthlim, dx, dy, self.dt,
                          comp_in=src[1], comp_out=whole[1], gin=gin, num_ghost=ng, aux=aux,
                          aux_bc_lower=alo, aux_bc_upper=ahi, whole=True, fast=self._fast)
            for a, b in zip(out, whole):
                if a is not None:
                    a.copy_(b[:, ng:-ng, ng:-ng])
            from .device import from_planar
            self.qbc = np.asfortranarray(from_planar(whole[0].double().cpu().numpy()))
        elif self.dimensional_split:
            self._work = K.step2_split(self._rp_struct, src[0], out[0], cfl_dev, lo, hi,
                                       self.reflect_index, self._method, self._mthlim, dx, dy,
                                       self.dt, comp_in=src[1], comp_out=out[1], gin=gin,
                                       num_ghost=self.num_ghost, work=self._work, aux=aux,
                                       aux_bc_lower=alo, aux_bc_upper=ahi, fast=self._fast)
        else:
            K.step2_unsplit(self._rp_struct, src[0], out[0], cfl_dev, lo, hi, self.reflect_index,
                            self._method, self._mthlim, dx, dy, self.dt, comp_in=src[1],
                            comp_out=out[1], gin=gin, num_ghost=self.num_ghost, aux=aux,
                            aux_bc_lower=alo, aux_bc_upper=ahi, fast=self._fast)
        cfl = float(cfl_dev.item())
        self.cfl.update_global_max(cfl)
        state.set_device_q(*out)
        self._scratch = qin

    # ---- device-resident loop
    def _device_loop_ok(self, solution):
        lo, hi = self._bc_codes()
        return (self.device_loop and not self.dimensional_split and self.before_step is None
                and self.step_source is None and BC.custom not in lo and BC.custom not in hi
                and solution.states[0].num_aux == 0)

    def evolve_to_time(self, solution, tend=None):
        if not self._is_set_up:
            self.setup(solution)
        if tend is None or not self._device_loop_ok(solution):
            return super(ClawSolver2D, self).evolve_to_time(solution, tend)
        tstart = solution.t
        max_steps = self.max_steps
        if not self.dt_variable:
            max_steps = int((tend - tstart + 1e-10) / self.dt)
            self.max_steps = max_steps
            if abs(max_steps * self.dt - (tend - tstart)) > 1e-5 * (tend - tstart):
                raise Exception("dt does not divide (tend-tstart) and dt is fixed!")
        if self.dt_variable == 1 and self.cfl_desired > self.cfl_max:
            raise Exception("Variable time-stepping and desired CFL > maximum CFL")
        if tend <= tstart:
            self.logger.info("Already at or beyond end time: no evolution required.")
            return self.status
        self.run_device(solution, tend, max_steps)
        return self.status

    def run_steps(self, solution, nsteps):
        """Attempt exactly nsteps steps (no end time) in the device loop."""
        if not self._is_set_up:
            self.setup(solution)
        if not self._device_loop_ok(solution):
            raise ValueError("run_steps needs the device loop: unsplit, no step callbacks, "
                             "built-in boundary conditions")
        return self.run_device(solution, float("inf"), int(nsteps), check_max_steps=False)

    def run_device(self, solution, tend, max_steps, check_max_steps=True):
        """Up to max_steps attempted steps towards tend, controlled on the device.
        Same arithmetic and decisions as Solver.evolve_to_time (solver.py:554-666)."""
        import torch
        from . import _lib as L
        state = solution.states[0]
        self.get_dt(solution.t, solution.t, tend, False)   # first attempt's dt (host)
        qin = state.device_q(self._dtype_code)
        other = self._take_scratch(qin)
        key = (qin[0].data_ptr(), other[0].data_ptr())
        if key in self._graphs or (other[0].data_ptr(), qin[0].data_ptr()) in self._graphs:
            # keep buffer identities fixed for captured graphs: slot 0 / slot 1
            k2 = key if key in self._graphs else (other[0].data_ptr(), qin[0].data_ptr())
            bufs = self._graphs[k2]["bufs"]
            cur = 0 if bufs[0][0].data_ptr() == qin[0].data_ptr() else 1
        else:
            bufs = (qin, other)
            cur = 0
        ctl = L.StepCtl()
        ctl.t, ctl.dt, ctl.tstart, ctl.tend = solution.t, self.dt, solution.t, tend
        ctl.cfl = self.cfl.get_cached_max()
        ctl.cfl_desired, ctl.cfl_max, ctl.dt_max = self.cfl_desired, self.cfl_max, self.dt_max
        ctl.cflmax, ctl.dtmin, ctl.dtmax = (self.status["cflmax"], self.status["dtmin"],
                                           self.status["dtmax"])
        ctl.dt_variable = int(bool(self.dt_variable))
        ctl.cur = cur
        ctl.max_steps = int(max_steps)
        raw = bytes(ctl)
        if self._ctl is None:
            self._ctl = torch.empty(len(raw), dtype=torch.uint8, device="cuda")
        self._ctl.copy_(torch.frombuffer(bytearray(raw), dtype=torch.uint8))
        grid = state.grid
        dx, dy = grid.delta
        lo, hi = self._bc_codes()
        args = (self._rp_struct, (bufs[0][0], bufs[1][0]), (bufs[0][1], bufs[1][1]), self._ctl,
                lo, hi, self.reflect_index, self._method, self._mthlim, dx, dy)
        launched = 0
        host = L.StepCtl()
        t_now, dt_now = solution.t, self.dt
        while launched < max_steps:
            # launch as many attempts as the remaining time needs (estimated
            # from the current dt) before reading the controller back: the
            # launches after it is done exit at once
            todo = _attempts_to_launch(max_steps - launched, tend, t_now, dt_now,
                                       self.graph_chunk)
            while todo > 0:
                n = min(self.graph_chunk, todo)
                if self.use_graph and n == self.graph_chunk:
                    self._replay_graph(bufs, args, n)
                else:
                    K.step2_unsplit_run(*args, n, fast=self._fast)
                launched += n
                todo -= n
            raw = self._ctl.cpu().numpy().tobytes()
            ctypes_memmove(host, raw)
            if host.done:
                break
            t_now, dt_now = host.t, host.dt
        solution.t = host.t
        self.dt = host.dt
        self.cfl.set_global_max(host.cfl)
        self.status["cflmax"], self.status["dtmin"], self.status["dtmax"] = (host.cflmax,
                                                                              host.dtmin, host.dtmax)
        self.status["numsteps"] += host.accepted
        state.set_device_q(*bufs[host.cur])
        self._scratch = bufs[1 - host.cur]
        self.accept_step = host.done != 2
        if host.done == 2:
            raise Exception("CFL too large, giving up!")
        if (check_max_steps and self.dt_variable and solution.t < tend
                and host.accepted == max_steps):
            raise Exception("Maximum number of timesteps have been taken")
        return host

    def _replay_graph(self, bufs, args, n):
        import torch
        key = (bufs[0][0].data_ptr(), bufs[1][0].data_ptr(), self._fast)
        ent = self._graphs.get(key)
        if ent is None:
            from .device import no_gc
            g = torch.cuda.CUDAGraph()
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            # no garbage collection while capturing: collecting an old graph
            # would destroy it mid-capture, which HIP refuses (an abort)
            with no_gc(), torch.cuda.stream(s):
                with torch.cuda.graph(g, stream=s):
                    K.step2_unsplit_run(*args, n, fast=self._fast)
            torch.cuda.current_stream().wait_stream(s)
            ent = {"graph": g, "bufs": bufs, "n": n}
            self._graphs[key] = ent
        ent["graph"].replay()


class ClawSolver3D(ClawSolver):
    """classic/solver.py:554-706: the dimension-split step (step3ds) and the
    unsplit step (step3, FP64, euler_3D)."""
    no_trans = 0
    trans_inc = 11
    trans_cor = 22

    def __init__(self, riemann_solver=None, claw_package=None):
        self.dimensional_split = True
        self.transverse_waves = self.trans_cor
        self.num_dim = 3
        self.reflect_index = [1, 2, 3]
        self.aux1 = None
        self.aux2 = None
        self.aux3 = None
        self.work = None
        self._buf3 = None
        super(ClawSolver3D, self).__init__(riemann_solver, claw_package)

    def setup(self, solution):
        if not self.dimensional_split:
            if self.precision != "fp64":
                raise NotImplementedError("the unsplit 3D step runs in FP64")
            if self.transverse_waves not in (0, 10, 11, 20, 21, 22):
                raise NotImplementedError("transverse_waves %r is not one of 0, 10, 11, 20, 21, "
                                          "22" % (self.transverse_waves,))
            if self.transverse_waves >= 20 and self.order != 2:
                raise NotImplementedError("the correction's transverse split needs order 2")
        super(ClawSolver3D, self).setup(solution)

    def step_hyperbolic(self, solution):
        state = solution.states[0]
        dx, dy, dz = state.grid.delta
        lo, hi = self._bc_codes()
        custom = BC.custom in lo or BC.custom in hi
        if custom and self.dimensional_split:
            raise NotImplementedError("custom boundary conditions in 3D: the unsplit step only")
        qin = state.device_q(self._dtype_code)
        out = self._take_scratch(qin)
        cfl_dev = self._cfl_buffer()
        K.cfl_reset(cfl_dev)
        if not self.dimensional_split:
            src, gin = qin, None
            if custom:
                src, gin = self._custom_bc_input(state, lo, hi)
                lo, hi = [4, 4, 4], [4, 4, 4]
            self._buf3 = K.step3_unsplit(self._rp_struct, src[0], out[0], cfl_dev, lo, hi,
                                         self.reflect_index, self._method, self._mthlim, dx, dy,
                                         dz, self.dt, num_ghost=self.num_ghost, gin=gin,
                                         work=self._buf3 if not isinstance(self._buf3, tuple)
                                         else None)
        else:
            self._buf3 = K.step3_split(self._rp_struct, qin[0], out[0], cfl_dev, lo, hi,
                                       self.reflect_index, self._method, self._mthlim, dx, dy,
                                       dz, self.dt, comp_in=qin[1], comp_out=out[1],
                                       work=self._buf3 if isinstance(self._buf3, tuple) else None,
                                       num_ghost=self.num_ghost, fast=self._fast)
        self.cfl.update_global_max(float(cfl_dev.item()))
        state.set_device_q(*out)
        self._scratch = qin


def ctypes_memmove(struct, raw):
    import ctypes
    ctypes.memmove(ctypes.addressof(struct), raw, ctypes.sizeof(struct))


def _attempts_to_launch(left, tend, t, dt, chunk):
    """Device-loop attempts to launch before the next read of the controller:
    all that are left for a fixed count (tend infinite), else the estimate of
    the attempts to tend at the current dt plus two, at least one chunk."""
    import math
    if not math.isfinite(tend) or not dt > 0:
        return left
    est = (tend - t) / dt
    if not math.isfinite(est):
        return min(left, chunk)
    return min(left, max(chunk, int(est) + 2))


def _fill_bcs_host(solver, state, qbc, lo, hi, aux=False):
    """Solver._apply_bcs on a host qbc (solver.py:351-519), for custom BCs; with
    aux=True the array is auxbc: the solver's aux functions, walls not negated."""
    ng = solver.num_ghost
    for idim, dim in enumerate(state.grid.dimensions):
        for upper, codes in ((False, lo), (True, hi)):
            b = codes[idim]
            if b == BC.custom:
                if aux:
                    fun = solver.user_aux_bc_upper if upper else solver.user_aux_bc_lower
                    fun(state, dim, state.t, None, qbc, ng)
                else:
                    fun = solver.user_bc_upper if upper else solver.user_bc_lower
                    fun(state, dim, state.t, qbc, solver.auxbc, ng)
                continue
            arr = np.moveaxis(qbc, idim + 1, 1)
            r = -1 if aux else solver.reflect_index[idim]
            for i in range(ng):
                if not upper:
                    if b == BC.extrap:
                        arr[:, i, ...] = arr[:, ng, ...]
                    elif b == BC.periodic:
                        arr[:, i, ...] = arr[:, -2 * ng + i, ...]
                    elif b == BC.wall:
                        arr[:, i, ...] = arr[:, 2 * ng - 1 - i, ...]
                        if r >= 0:
                            arr[r, i, ...] = -arr[r, 2 * ng - 1 - i, ...]
                else:
                    if b == BC.extrap:
                        arr[:, -i - 1, ...] = arr[:, -ng - 1, ...]
                    elif b == BC.periodic:
                        arr[:, -ng + i, ...] = arr[:, ng + i, ...]
                    elif b == BC.wall:
                        arr[:, -i - 1, ...] = arr[:, -2 * ng + i, ...]
                        if r >= 0:
                            arr[r, -i - 1, ...] = -arr[r, -2 * ng + i, ...]
