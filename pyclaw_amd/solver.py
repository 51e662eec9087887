"""Solver base: time-step control, accept/reject, boundary-condition codes.

Restates pyclaw/solver.py:35-666 (Solver.evolve_to_time, get_dt, get_dt_new,
accept_reject_step, is_valid) for solvers whose state lives on the GPU.  The
step's backup copy (solver.py:610-612) is not a copy here: the kernels write a
new buffer, so the pre-step buffers are the backup and a rejected step only
restores the tensor references.
"""
import logging

import numpy as np

from .cfl import CFL


class CFLError(Exception):
    def __init__(self, msg):
        super(CFLError, self).__init__(msg)


class BC(object):
    """solver.py:13-19"""
    custom = 0
    extrap = 1
    periodic = 2
    wall = 3


def default_compute_gauge_values(q, aux):
    return q


class Solver(object):
    """Common attributes and the evolve_to_time loop of every solver."""

    def __setattr__(self, key, value):
        if not hasattr(self, "_isinitialized"):
            self.__dict__["_isinitialized"] = False
        if self._isinitialized and not hasattr(self, key):
            raise TypeError("%s has no attribute %s" % (self.__class__, key))
        object.__setattr__(self, key, value)

    @property
    def all_bcs(self):
        return self.bc_lower, self.bc_upper

    @all_bcs.setter
    def all_bcs(self, all_bcs):
        for i in range(self.num_dim):
            self.bc_lower[i] = all_bcs
            self.bc_upper[i] = all_bcs

    def __init__(self, riemann_solver=None, claw_package=None):
        self.logger = logging.getLogger("pyclaw.solver")
        self.dt_initial = 0.1
        self.dt_max = 1e99
        self.max_steps = 10000
        self.dt_variable = True
        self.num_waves = None
        self.qbc = None
        self.auxbc = None
        self.rp = None
        self.fmod = None
        self._is_set_up = False
        self._use_old_bc_sig = False
        self.accept_step = True
        self.before_step = None
        self.dt = self.dt_initial
        self.cfl = CFL(self.cfl_desired)
        self.status = {"cflmax": -np.inf, "dtmin": np.inf, "dtmax": -np.inf, "numsteps": 0}
        self.bc_lower = [None] * self.num_dim
        self.bc_upper = [None] * self.num_dim
        self.aux_bc_lower = [None] * self.num_dim
        self.aux_bc_upper = [None] * self.num_dim
        self.user_bc_lower = None
        self.user_bc_upper = None
        self.user_aux_bc_lower = None
        self.user_aux_bc_upper = None
        self.num_eqn = None
        self.num_waves = None
        self.fwave = None
        # Arithmetic of the device state: "fp64" (bitwise the reference's step),
        # "fp64fast" (FP64 with a*b + c contracted to FMAs: within about 1e-13
        # relative of "fp64" per step) or "fp32k" (FP32 state with
        # Kahan-compensated updates, FP64 CFL).
        self.precision = "fp64"
        self.compute_gauge_values = default_compute_gauge_values
        self._cfl_dev = None
        self._pinned = None
        self._scratch = None
        if riemann_solver is not None:
            self.rp = riemann_solver
            from . import riemann
            rp_name = riemann_solver.__name__.split(".")[-1]
            self.num_eqn = riemann.static.num_eqn.get(rp_name, None)
            self.num_waves = riemann.static.num_waves.get(rp_name, None)
            self.fwave = riemann.static.fwave.get(rp_name, None)
        self._isinitialized = True

    # ---- validation (solver.py:247-287)
    def is_valid(self):
        valid = True
        reason = None
        if any([m == BC.custom for m in self.bc_lower]) and self.user_bc_lower is None:
            valid, reason = False, "Lower custom BC function has not been set."
        if any([m == BC.custom for m in self.bc_upper]) and self.user_bc_upper is None:
            valid, reason = False, "Upper custom BC function has not been set."
        if self.num_waves is None:
            valid, reason = False, "solver.num_waves has not been set."
        if self.num_eqn is None:
            valid, reason = False, "solver.num_eqn has not been set."
        if self.fwave is None:
            valid, reason = False, "solver.fwave has not been set."
        if (None in self.bc_lower) or (None in self.bc_upper):
            valid, reason = False, "One of the boundary conditions has not been set."
        if reason is not None:
            self.logger.debug(reason)
        return valid, reason

    def setup(self, solution):
        self._is_set_up = True

    # ---- device helpers
    @property
    def _dtype_code(self):
        """The storage type of the device state: 0 FP64 ("fp64", "fp64fast"), 1
        FP32 + Kahan ("fp32k")."""
        if self.precision in ("fp64", "fp64fast"):
            return 0
        if self.precision == "fp32k":
            return 1
        raise ValueError("solver.precision must be 'fp64', 'fp64fast' or 'fp32k'")

    @property
    def _fast(self):
        """FP64 steps in the FMA-contracted variant (precision "fp64fast")."""
        return self.precision == "fp64fast"

    def _cfl_buffer(self):
        import torch
        if self._cfl_dev is None:
            self._cfl_dev = torch.zeros(1, dtype=torch.float64, device="cuda")
        return self._cfl_dev

    def _read_back(self, t):
        """Host list of a small device tensor: one copy into pinned host memory on
        the current stream, then a wait for that stream."""
        import torch
        if self._pinned is None:
            self._pinned = {}
        pinned = self._pinned
        key = (t.numel(), t.dtype)
        h = pinned.get(key)
        if h is None:
            h = pinned[key] = torch.empty(t.numel(), dtype=t.dtype, pin_memory=True)
        h.copy_(t.reshape(-1), non_blocking=True)
        torch.cuda.current_stream().synchronize()
        return h.tolist()

    def _bc_codes(self):
        """Kernel boundary codes; raises for unset BCs as _bc_lower does (solver.py:474-478)."""
        lo, hi = [], []
        for d in range(self.num_dim):
            for side, codes, out in (("Lower", self.bc_lower, lo), ("Upper", self.bc_upper, hi)):
                b = codes[d]
                if b is None:
                    raise Exception("%s boundary condition not specified for either q or aux." % side)
                if b not in (BC.custom, BC.extrap, BC.periodic, BC.wall):
                    raise NotImplementedError("Boundary condition %s not implemented" % b)
                if b == BC.wall and self.reflect_index[d] >= self.num_eqn:
                    raise IndexError("wall boundary: reflect_index %d >= num_eqn %d"
                                     % (self.reflect_index[d], self.num_eqn))
                out.append(int(b))
        return lo, hi

    def _aux_bc_codes(self):
        """Kernel codes of solver.aux_bc_lower/upper (extrap, periodic, wall: a wall
        mirrors aux without negation, solver.py:439-519; custom: the solver fills
        the ghosts on the host)."""
        lo, hi = [], []
        for d in range(self.num_dim):
            for side, codes, out in (("Lower", self.aux_bc_lower, lo),
                                     ("Upper", self.aux_bc_upper, hi)):
                b = codes[d]
                if b is None:
                    raise Exception("%s boundary condition not specified for either q or aux." % side)
                if b not in (BC.custom, BC.extrap, BC.periodic, BC.wall):
                    raise NotImplementedError("Boundary condition %s not implemented" % b)
                if b == BC.custom and (self.user_aux_bc_upper if side == "Upper"
                                       else self.user_aux_bc_lower) is None:
                    raise Exception("%s custom aux BC function has not been set." % side)
                out.append(int(b))
        return lo, hi

    def _aux_input(self, state):
        """Device aux and its boundary codes.  With a custom aux BC the ghosts are
        filled on the host in the reference's order (solver.py:371-437; the user
        function gets qbc = None: q lives on the device) and stored (code 4)."""
        from .device import q_to_device
        alo, ahi = self._aux_bc_codes()
        if BC.custom not in alo and BC.custom not in ahi:
            return state.device_aux(self._dtype_code), alo, ahi
        auxbc = self._host_auxbc(state)
        return q_to_device(auxbc, self._dtype_code), [4] * self.num_dim, [4] * self.num_dim

    def _host_auxbc(self, state):
        """solver.auxbc on the host: aux with its ghosts filled by the aux BCs in
        the reference's order (walls mirror without negation)."""
        alo, ahi = self._aux_bc_codes()
        ng = self.num_ghost
        aux = np.asarray(state.aux)
        auxbc = np.zeros((aux.shape[0],) + tuple(n + 2 * ng for n in aux.shape[1:]), order="F")
        auxbc[(slice(None),) + (slice(ng, -ng),) * self.num_dim] = aux
        self.auxbc = auxbc
        from .classic import _fill_bcs_host
        _fill_bcs_host(self, state, auxbc, alo, ahi, aux=True)
        return auxbc

    # ---- time-step control (solver.py:524-552)
    def accept_reject_step(self, state):
        cfl = self.cfl.get_cached_max()
        if cfl > self.cfl_max:
            return False
        return True

    def get_dt_new(self):
        cfl = self.cfl.get_cached_max()
        self.dt = min(self.dt_max, self.dt * self.cfl_desired / cfl)

    def get_dt(self, t, tstart, tend, take_one_step):
        cfl = self.cfl.get_cached_max()
        if self.dt_variable and self.dt_old is not None:
            if cfl > 0.0:
                self.get_dt_new()
                self.status["dtmin"] = min(self.dt, self.status["dtmin"])
                self.status["dtmax"] = max(self.dt, self.status["dtmax"])
            else:
                self.dt = self.dt_max
        else:
            self.dt_old = self.dt
        if not take_one_step:
            if t + self.dt > tend and tstart < tend:
                self.dt = tend - t
            if tend - t - self.dt < 1.e-14 * t:
                self.dt = tend - t

    # ---- the stepping loop (solver.py:554-666)
    def evolve_to_time(self, solution, tend=None):
        if not self._is_set_up:
            self.setup(solution)
        take_one_step = tend is None
        tstart = solution.t
        num_steps = 0
        if not self.dt_variable:
            if take_one_step:
                self.max_steps = 1
            else:
                self.max_steps = int((tend - tstart + 1e-10) / self.dt)
                if abs(self.max_steps * self.dt - (tend - tstart)) > 1e-5 * (tend - tstart):
                    raise Exception("dt does not divide (tend-tstart) and dt is fixed!")
        if self.dt_variable == 1 and self.cfl_desired > self.cfl_max:
            raise Exception("Variable time-stepping and desired CFL > maximum CFL")
        if not take_one_step and tend <= tstart:
            self.logger.info("Already at or beyond end time: no evolution required.")
            self.max_steps = 0

        for n in range(self.max_steps):
            state = solution.state
            if self.dt_variable:
                backup = self._backup(state)
                told = solution.t
            if self.before_step is not None:
                self.before_step(self, solution.states[0])
            self.step(solution, take_one_step, tstart, tend)
            cfl = self.cfl.get_cached_max()
            self.accept_step = self.accept_reject_step(state)
            if self.accept_step:
                self.status["cflmax"] = max(cfl, self.status["cflmax"])
                if self.dt_variable:
                    solution.t += self.dt
                else:
                    solution.t = tstart + (n + 1) * self.dt
                self.logger.debug("Step %i  CFL = %f   dt = %f   t = %f" % (n, cfl, self.dt, solution.t))
                num_steps += 1
                self.status["numsteps"] += 1
            else:
                self.logger.debug("Rejecting time step, CFL number too large")
                if self.dt_variable:
                    self._restore(state, backup)
                    solution.t = told
                else:
                    self.status["cflmax"] = max(cfl, self.status["cflmax"])
                    raise Exception("CFL too large, giving up!")
            if take_one_step:
                break
            elif solution.t >= tend:
                break
        if not take_one_step:
            if self.dt_variable and solution.t < tend and num_steps == self.max_steps:
                raise Exception("Maximum number of timesteps have been taken")
        return self.status

    def _backup(self, state):
        """The pre-step device tensors (kept alive, never overwritten by the step)."""
        dev = state.device_q(self._dtype_code)
        return dev

    def _restore(self, state, backup):
        rejected = state._dev
        state.set_device_q(*backup)
        if rejected is not None and rejected[0] is not backup[0]:
            self._scratch = rejected

    def _take_scratch(self, like):
        """Output buffers for a step: never the tensors the state holds."""
        import torch
        sc = self._scratch
        if sc is None or sc[0].shape != like[0].shape or sc[0].dtype != like[0].dtype \
                or sc[0] is like[0]:
            q = torch.empty_like(like[0])
            comp = torch.empty_like(like[1]) if like[1] is not None else None
            sc = (q, comp)
        self._scratch = None
        return sc

    def step(self, solution, take_one_step, tstart, tend):
        raise NotImplementedError("No stepping routine has been defined!")

    def write_gauge_values(self, solution):
        pass

    def __str__(self):
        output = "Solver Status:\n"
        for (k, v) in self.status.items():
            output = "\n".join((output, "%s = %s" % (k.rjust(25), v)))
        return output
