"""SharpClaw 2D solver (pyclaw/sharpclaw/solver.py:28-983) on the GPU.

The semi-discrete operator dq = dt*L(q) (ghost fill + WENO5 reconstruction +
Riemann solves) runs in clawhip_sharp2_stage, fused with the Runge-Kutta stage
combination; Euler, SSP22, SSP33, SSP104, a general explicit Runge-Kutta method
(RK: solver.a, b, c), the multistep SSPLMMk2/k3 and a general linear multistep
method (LMM: solver.alpha, beta; fixed dt) keep the reference's stage sequence,
step-size control and arithmetic (sharpclaw/solver.py:283-490, 595-690), so the
FP64 path equals the reference step bit for bit.  precision = "fp32k" runs SSP33 with FP32 stages and a
Kahan-compensated final update.
"""
import numpy as np
import torch

from . import kernels as K
from .classic import _attempts_to_launch
from .solver import BC, Solver


def ctypes_sizeof_sharpctl():
    import ctypes
    from . import _lib as L
    return ctypes.sizeof(L.SharpCtl)


class _StageState(object):
    """The state one operator evaluation sees (the reference's register States,
    sharpclaw/solver.py:306-430): the solver's state with q replaced by the
    stage's (a host array) and t by the stage's time (None: the state's own),
    everything else read from the state."""

    def __init__(self, state, q, t=None):
        self._state = state
        self.q = q
        if t is not None:
            self.t = t

    def __getattr__(self, name):
        return getattr(self._state, name)


class SharpClawSolver(Solver):
    _sspcoeff = {"Euler": 1.0, "SSP22": 1.0, "SSP33": 1.0, "SSP104": 6.0,
                 "SSPLMM32": 0.5, "SSPLMM43": 1. / 3., "SSPLMM53": 0.5, "RK": None,
                 "LMM": None}
    _cfl_default = {"SSP104": [2.45, 2.5], "SSPLMM32": [0.24, 0.25],
                    "SSPLMM43": [0.15, 1. / 6.], "SSPLMM53": [0.24, 0.25]}
    _integrators = ("Euler", "SSP22", "SSP33", "SSP104", "SSPLMMk2", "SSPLMMk3", "RK", "LMM")

    def __init__(self, riemann_solver=None, claw_package=None):
        self.limiters = [1]
        self.lim_type = 2
        self.weno_order = 5
        self.time_integrator = "SSP104"
        self.char_decomp = 0
        self.tfluct_solver = False
        self.tfluct = None
        self.aux_time_dep = False
        self.kernel_language = "Fortran"
        self.num_ghost = 3
        self.fwave = False
        self.cfl_desired = None
        self.cfl_max = None
        self.dq_src = None
        self.call_before_step_each_stage = False
        self._mthlim = self.limiters
        self._method = None
        self._registers = None
        self.dq_dt = None
        self.dt_old = None
        self._rp_struct = None
        self._bufs = {}
        self.indicator = None
        self._flags = None
        self._qaux = None     # (device aux, lower codes, upper codes) for the current step
        self._state = None    # the state being stepped
        self._t0 = 0.0        # the start time of the step being taken
        # general Runge-Kutta coefficients (time_integrator 'RK')
        self.a = None
        self.b = None
        self.c = None
        # multistep methods (sharpclaw/solver.py:186-195)
        self.alpha = None
        self.beta = None
        self.sspcoeff0 = None
        self.lmm_steps = 4
        self.sspcoeff = None
        self.prev_dq_dt_values = []
        self.prev_dt_values = []
        self.prev_dtFE_values = []
        self.check_lmm_cond = False
        self.lmm_cond = True
        # evolve_to_time(solution, tend) and run_steps run SSP33 as device-
        # controlled launches (clawhip_sharp2_ssp33_run, graph-captured chunks of
        # graph_chunk attempts) when no host callback takes part
        self.device_loop = True
        self.use_graph = True
        self.graph_chunk = 16
        self._graphs = {}
        self._ctl = None
        self._stage_events = None   # time_stages: events of the eager launches
        self._stage_ev_pos = 0
        self._stage_ran = []        # time_stages: whether each timed stage's gate was open
        super(SharpClawSolver, self).__init__(riemann_solver, claw_package)

    # ---- sharpclaw/solver.py:202-265
    def setup(self, solution):
        if self.lim_type == 2:
            self.num_ghost = (self.weno_order + 1) // 2
        if self.lim_type != 2 or self.weno_order != 5:
            raise NotImplementedError("pyclaw_amd implements lim_type=2 with weno_order=5")
        if self.char_decomp not in (0, 2):
            raise NotImplementedError("char_decomp must be 0 (component) or 2 (characteristic)")
        if self.tfluct_solver or self.fwave:
            raise NotImplementedError("tfluct solvers and f-waves are not supported")
        if self.call_before_step_each_stage and self.before_step is not None:
            raise NotImplementedError("call_before_step_each_stage is not supported: "
                                      "before_step runs once per step")
        custom = BC.custom in list(self.bc_lower) + list(self.bc_upper)
        if self.indicator is not None and (self.dq_src is not None or custom):
            raise NotImplementedError("the indicator runs without dq_src and custom BCs")
        if self.time_integrator not in self._integrators:
            raise NotImplementedError("time_integrator must be one of %s" % (self._integrators,))
        if self.precision == "fp32k" and self.time_integrator != "SSP33":
            raise NotImplementedError("the FP32 path implements SSP33")
        if self.indicator is not None and self.time_integrator not in ("SSP33", "SSP104"):
            raise NotImplementedError("the indicator runs with SSP33 or SSP104")
        if self.time_integrator == "LMM":
            assert self.dt_variable == False, \
                "Must set solver.dt_variable=False for LMM integrator."
            if self.alpha is None or self.beta is None:
                raise ValueError("the LMM integrator needs solver.alpha and solver.beta")
        if self.time_integrator == "RK" and (self.a is None or self.b is None):
            raise ValueError("the RK integrator needs solver.a and solver.b")
        if self.check_lmm_cond:
            raise NotImplementedError("check_lmm_cond (the third-order SSPLMM a-posteriori "
                                      "conditions) is not supported")
        try:
            if "SSPLMM" in self.time_integrator:
                if self.time_integrator == "SSPLMMk2":
                    assert 3 <= self.lmm_steps, \
                        "Must set solver.lmm_steps greater than 2 for 2nd order SSPLMM integrator."
                    self.sspcoeff = (self.lmm_steps - 2.) / (self.lmm_steps - 1.)
                    if self.cfl_max is None:
                        self.cfl_desired = 0.14 * self.sspcoeff
                        self.cfl_max = 0.15 * self.sspcoeff
                else:
                    assert 4 <= self.lmm_steps <= 5, \
                        "Must set solver.lmm_steps equal to 4 or 5 for 3rd order SSPLMM integrator."
                    self.sspcoeff = (self.lmm_steps - 3.) / (self.lmm_steps - 1.)
                    if self.cfl_max is None:
                        self.cfl_desired = 0.48 * self.sspcoeff
                        self.cfl_max = 0.5 * self.sspcoeff
            elif self.cfl_max is None:
                self.cfl_desired = self._cfl_default[self.time_integrator][0]
                self.cfl_max = self._cfl_default[self.time_integrator][1]
            if self.cfl_desired is None:
                self.cfl_desired = 0.9 * self.cfl_max
        except KeyError:
            raise KeyError("Maximum CFL number is not provided.")
        if self.time_integrator == "RK" and self.c is None:
            # the stage times are t + dt*c[i] (sharpclaw/solver.py:338)
            raise ValueError("the RK integrator needs solver.c (its stage times)")
        self._registers = None
        self.prev_dq_dt_values, self.prev_dt_values, self.prev_dtFE_values = [], [], []
        self._set_mthlim()
        state = solution.states[0]
        if self.rp is None or not hasattr(self.rp, "kind"):
            raise Exception("the solver needs one of the pyclaw_amd.riemann modules")
        rp_aux = getattr(self.rp, "num_aux", 0)
        if state.index_capa >= 0:
            raise NotImplementedError("capacity functions are not supported by SharpClaw on the device")
        if rp_aux and not rp_aux <= state.num_aux <= 3:
            raise NotImplementedError("the Riemann solver reads %d aux fields; the device "
                                      "operator takes %d to 3, the state has %d"
                                      % (rp_aux, rp_aux, state.num_aux))
        if rp_aux and (self.char_decomp != 0 or self.indicator is not None):
            raise NotImplementedError("solvers with aux fields run with char_decomp = 0 and "
                                      "no indicator")
        self._rp_struct = K.make_rp(self.rp.kind, state.problem_data, meqn=state.num_eqn)
        self._is_set_up = True

    def _reads_aux(self, state):
        # aux fields only the user's callbacks read stay on the host
        return state.num_aux > 0 and getattr(self.rp, "num_aux", 0) > 0

    def _set_mthlim(self):
        self._mthlim = self.limiters
        if not isinstance(self.limiters, list):
            self._mthlim = [self._mthlim]
        n = self.num_eqn if self.char_decomp == 0 else self.num_waves
        if len(self._mthlim) == 1:
            self._mthlim = self._mthlim * n
        if len(self._mthlim) != n:
            raise Exception("Length of solver.limiters is not equal to 1 or to solver.num_eqn/num_waves")

    # ---- sharpclaw/solver.py:653-685
    def get_dt_new(self):
        self.dt_old = self.dt
        cfl = self.cfl.get_cached_max()
        if "SSPLMM" in self.time_integrator:
            step_index = self.status["numsteps"] + 1
            if step_index < self.lmm_steps:
                # step size of the starting method
                sspcoeff_ratio = self.sspcoeff0 / self.sspcoeff
                self.dt = sspcoeff_ratio * self.dt * self.cfl_desired / cfl
            elif self.accept_step:
                # this step size keeps the multistep method within its CFL bound
                s = self.lmm_steps
                p = int(self.time_integrator[-1])
                mu = min([self.prev_dtFE_values[i] for i in range(s)])
                H = sum(self.prev_dt_values[1:])
                self.dt = H * mu / (H + (p - 1) * mu)
        else:
            self.dt = self.dt * self.cfl_desired / cfl

    def _num_registers(self):
        """The number of saved states (sharpclaw/solver.py:596-610)."""
        if self.time_integrator == "RK":
            return len(self.b) + 1
        if self.time_integrator == "LMM":
            return len(self.alpha)
        return self.lmm_steps

    # ---- sharpclaw/solver.py:622-650
    def accept_reject_step(self, state):
        cfl = self.cfl.get_cached_max()
        if "LMM" in self.time_integrator:
            step_index = self.status["numsteps"] + 1
            if step_index < self._num_registers():   # only the starting steps are checked
                ratio = 1. if self.time_integrator == "LMM" else self.sspcoeff0 / self.sspcoeff
                return not (cfl > ratio * self.cfl_max)
            return True
        return not (cfl > self.cfl_max)

    def _buf(self, name, like):
        t = self._bufs.get(name)
        if t is None or t.shape != like.shape or t.dtype != like.dtype:
            t = torch.empty_like(like)
            self._bufs[name] = t
        return t

    def _stage(self, state_q, out, P, cfl=None, stage_t=None, **kw):
        """One operator evaluation (dq) fused with a stage.  cfl=None: the stage's
        Courant number goes into self.cfl at once; otherwise it is max-reduced
        into the given device buffer, which the caller reads later.  stage_t:
        the time of the stage's state, which host callbacks see (None: the
        step's start time)."""
        lo, hi = self._bc_codes()
        if self.dq_src is not None or BC.custom in lo or BC.custom in hi:
            P, lo, hi = self._host_stage_input(P, kw, lo, hi, stage_t)
        dx, dy = self._grid.delta
        now = cfl is None
        if now:
            cfl = self._cfl_buffer()
            K.cfl_reset(cfl)
        qaux, alo, ahi = self._qaux if self._qaux is not None else (None, None, None)
        K.sharp2_stage(self._rp_struct, self.char_decomp, out, cfl, lo, hi, self.reflect_index,
                       self.dt, dx, dy, P, flags=self._flags,
                       flag_list=self.indicator.flag_list if self._flags is not None else None,
                       recon=self.indicator.recon_buffer(out) if self._flags is not None
                       else None, qaux=qaux, qaux_bc_lower=alo, qaux_bc_upper=ahi,
                       fast=self._fast, **kw)
        if now:
            self.cfl.update_global_max(self._read_back(cfl)[0])

    def _host_stage_input(self, P, kw, lo, hi, stage_t=None):
        """The stage input on the host, for what runs there: the user's custom BC
        functions fill its ghosts (solver.py:351-437, on the stage's state, whose
        time is stage_t) and dq_src adds its increment
        (sharpclaw/solver.py:537-540).  Returns the device input (with stored
        ghosts under custom BCs) and its BC codes; kw gets the source and loses
        the in-kernel stage-input form (y_mode, Q).  kw keeps s: the FP32
        stage-2 combination (OUT_F32_S2) reads it as the step's dt."""
        from .classic import _fill_bcs_host
        from .device import q_from_device, q_to_device
        y_mode, Q, s = kw.pop("y_mode", K.Y_ARRAY), kw.pop("Q", None), kw.get("s", 0.0)
        y = q_from_device(P)
        if y_mode == K.Y_AXPY:
            y = y + s * q_from_device(Q)
        elif y_mode == K.Y_AXPY6:
            y = y + (s * q_from_device(Q)) / 6.
        st = _StageState(self._state, y, stage_t)
        if self.dq_src is not None:
            S = np.asarray(self.dq_src(self, st, self.dt))
            kw["src"] = q_to_device(S, self._dtype_code)
        if BC.custom not in lo and BC.custom not in hi:
            return q_to_device(y, self._dtype_code), lo, hi
        ng = self.num_ghost
        qbc = np.zeros((y.shape[0],) + tuple(n + 2 * ng for n in y.shape[1:]), order="F")
        qbc[(slice(None),) + (slice(ng, -ng),) * 2] = y
        if self._state.num_aux > 0:
            self._host_auxbc(self._state)
        _fill_bcs_host(self, st, qbc, lo, hi)
        return q_to_device(qbc, self._dtype_code), [4, 4], [4, 4]

    def _step_buffer(self, meqn):
        """Device float64 [Courant number of the later stages, totals of the
        meqn fields, non-positive cell count], read back with one copy."""
        t = self._bufs.get("step_buf")
        if t is None or t.numel() != meqn + 2:
            t = torch.zeros(meqn + 2, dtype=torch.float64, device="cuda")
            self._bufs["step_buf"] = t
        return t

    def _stage_cfl(self):
        """Device buffer for the Courant number of a step's later stages."""
        t = self._bufs.get("cfl_stages")
        if t is None:
            t = torch.zeros(1, dtype=torch.float64, device="cuda")
            self._bufs["cfl_stages"] = t
        K.cfl_reset(t)
        return t

    # ---- sharpclaw/solver.py:283-430
    def step(self, solution, take_one_step, tstart, tend):
        state = solution.states[0]
        self._grid = state.grid
        self._state = state
        q, comp = state.device_q(self._dtype_code)
        # aux and its ghosts once per step (dq_hyperbolic applies the aux BCs of
        # the step's state at every stage, sharpclaw/solver.py:848-898)
        self._qaux = self._aux_input(state) if self._reads_aux(state) else None
        self._flags = None
        if self.indicator is not None:
            self._flags = self.indicator.step_flags(q, self._bc_codes(), self.reflect_index,
                                                    self.cfl_max)
        sb = self._step_buffer(q.shape[0])
        cs = sb[0:1]
        K.cfl_reset(cs)   # before stage 1: off the path between its CFL read and stage 2
        step_index = self.status["numsteps"] + 1
        if self.accept_step:
            self.cfl.set_global_max(0.)
            dqdt = self._buf("dq_dt_new", q)
            self._stage(q, dqdt, q, out_mode=K.OUT_DQDT)
            # keep the previous dq_dt buffer free for the next evaluation
            self._bufs["dq_dt_new"], self._bufs["dq_dt"] = self._bufs.get("dq_dt"), dqdt
            self.dq_dt = dqdt
        if "LMM" in self.time_integrator:
            step_index = self._update_saved_values(q, state, step_index)
        self.get_dt(solution.t, tstart, tend, take_one_step)
        cfl = self.cfl.get_cached_max()
        self.cfl.set_global_max(self.dt / self.dt_old * cfl)
        self.dt_old = self.dt
        out_q = self._take_scratch((q, comp))
        # the later stages max-reduce their Courant numbers on the device; one
        # read at the end of the step (max is order-free, so self.cfl ends as the
        # stage-by-stage updates of sharpclaw/solver.py:283-430 leave it)
        self._t0 = solution.t
        if "LMM" in self.time_integrator:
            self._lmm_step(q, comp, out_q, step_index, cs)
        else:
            self._rk_stages(q, comp, self.dq_dt, out_q, self.dt, cs)
        if self._flags is not None:
            out_q = self._judge(q, comp, out_q, sb)
        else:
            self.cfl.update_global_max(self._read_back(cs)[0])
        state.set_device_q(*out_q)
        self._scratch = (q, comp)
        self._flags = None
        return True

    def _update_saved_values(self, q, state, step_index):
        """sharpclaw/solver.py:432-490: the saved states, step sizes and
        forward-Euler step sizes of the multistep method; the oldest saved state
        is rolled to the end and overwritten with the current one."""
        nreg = self._num_registers()
        if self._registers is None:
            self._registers = [torch.empty_like(q) for _ in range(nreg)]
        if self.prev_dt_values == []:
            # the very first step
            self._registers[-1].copy_(q)
            self.prev_dq_dt_values.append(self.dq_dt.clone())
            self.prev_dt_values.append(0.)   # not used
            if "SSPLMM" in self.time_integrator:
                cfl = self.cfl.get_cached_max()
                dtFE = self.dt / cfl * self.cfl_max / self.sspcoeff
                self.prev_dtFE_values.append(dtFE)
                self.sspcoeff0 = self._sspcoeff["SSP22"]
        elif self.accept_step and self.time_integrator == "LMM":
            if step_index <= nreg:
                self.prev_dq_dt_values.append(self.dq_dt.clone())
            else:
                self.prev_dq_dt_values = self.prev_dq_dt_values[1:] + self.prev_dq_dt_values[:1]
                self.prev_dq_dt_values[-1].copy_(self.dq_dt)
            self._registers = self._registers[1:] + self._registers[:1]
            self._registers[-1].copy_(q)
        elif self.accept_step:
            cfl = self.cfl.get_cached_max()
            dtFE = self.dt / cfl * self.cfl_max / self.sspcoeff
            if step_index <= self.lmm_steps:   # start-up
                if self.time_integrator == "SSPLMMk3":
                    self.prev_dq_dt_values.append(self.dq_dt.clone())
                self.prev_dt_values.append(self.dt_old)
                self.prev_dtFE_values.append(dtFE)
            else:
                if self.time_integrator == "SSPLMMk3":
                    self.prev_dq_dt_values = self.prev_dq_dt_values[1:] + self.prev_dq_dt_values[:1]
                    self.prev_dq_dt_values[-1].copy_(self.dq_dt)
                self.prev_dt_values = self.prev_dt_values[1:] + self.prev_dt_values[:1]
                self.prev_dt_values[-1] = self.dt_old
                self.prev_dtFE_values = self.prev_dtFE_values[1:] + self.prev_dtFE_values[:1]
                self.prev_dtFE_values[-1] = dtFE
            self._registers = self._registers[1:] + self._registers[:1]
            self._registers[-1].copy_(q)
        return step_index

    def _lmm_step(self, q, comp, out_q, step_index, cs):
        """sharpclaw/solver.py:351-384: SSP22 (SSPLMM) or SSP104 (LMM) for the
        starting values, then the multistep combination (no operator evaluation
        beyond dq_dt)."""
        dt = self.dt
        if self.time_integrator == "LMM":
            regs, prev = self._registers, self.prev_dq_dt_values
            if step_index < len(regs):
                self._rk_stages(q, comp, self.dq_dt, out_q, dt, cs, method="SSP104")
                return
            al, be = self.alpha, self.beta
            self._lmm(4, regs[-1], regs[-1], (0.0, al[-1], be[-1] * dt), out_q[0], v=prev[-1])
            for i in range(self.lmm_steps - 1):
                self._lmm(5, out_q[0], regs[i], (0.0, al[i], be[i] * dt), out_q[0], v=prev[i])
            return
        num_steps = self.lmm_steps
        if step_index < num_steps:
            self._stage(q, out_q[0], q, cfl=cs, stage_t=self._t0 + dt, Q=self.dq_dt, s=dt,
                        y_mode=K.Y_AXPY, out_mode=K.OUT_SSP22, A=q)
            return
        omega_k_minus_1 = sum(self.prev_dt_values[1:]) / dt
        if self.time_integrator == "SSPLMMk2":
            r = (omega_k_minus_1 - 1.) / omega_k_minus_1
            delta = 1. / omega_k_minus_1 ** 2
            beta = (omega_k_minus_1 + 1.) / omega_k_minus_1
            self._lmm(1, q, self.dq_dt, (beta, r, dt, delta), out_q[0],
                         R=self._registers[-num_steps])
        else:
            omega_k = omega_k_minus_1 + 1.
            r = (omega_k_minus_1 - 2.) / omega_k_minus_1
            delta0 = (4 * omega_k - omega_k_minus_1 ** 2) / omega_k_minus_1 ** 3
            beta0 = omega_k / omega_k_minus_1 ** 2
            beta_k_minus_1 = omega_k ** 2 / omega_k_minus_1 ** 2
            self._lmm(2, q, self.prev_dq_dt_values[-1],
                         (beta_k_minus_1, r, dt, r * beta0 + delta0, beta0 * dt), out_q[0],
                         R=self._registers[-num_steps], v=self.prev_dq_dt_values[-num_steps])

    def _rk_stages(self, q, comp, dq_dt, out_q, dt, cs, method=None):
        """The Runge-Kutta stages after the first operator evaluation
        (sharpclaw/solver.py:306-348, 388-430), into out_q.  Each stage's state
        carries the reference's stage time (the registers' t), computed from the
        step's start time self._t0 with the reference's arithmetic."""
        ti = self.time_integrator if method is None else method
        t0 = self._t0
        if ti == "Euler":
            self._lmm(0, q, dq_dt, (0.0, dt), out_q[0])
        elif ti == "SSP22":
            self._stage(q, out_q[0], q, cfl=cs, stage_t=t0 + dt, Q=dq_dt, s=dt,
                        y_mode=K.Y_AXPY, out_mode=K.OUT_SSP22, A=q)
        elif ti == "RK":
            # registers[0] = dt*dq_dt; stage i: q + sum_j a[i,j]*registers[j], term
            # by term, then registers[i] = dq(stage i); q += b[j]*registers[j]
            a, b = self.a, self.b
            regs = [self._buf("rk%d" % j, q) for j in range(len(b))]
            self._lmm(3, dq_dt, dq_dt, (0.0, dt), regs[0])
            y = self._buf("rk_y", q)
            for i in range(1, len(b)):
                y.copy_(q)
                for j in range(i):
                    self._lmm(0, y, regs[j], (0.0, float(a[i][j])), y)
                self._stage(q, regs[i], y, cfl=cs, stage_t=t0 + dt * self.c[i],
                            out_mode=K.OUT_D)
            out_q[0].copy_(q)
            for j in range(len(b)):
                self._lmm(0, out_q[0], regs[j], (0.0, float(b[j])), out_q[0])
        elif ti == "SSP33":
            r2 = self._buf("r2", q)
            # register times t + dt, then t + 0.5*dt (sharpclaw/solver.py:313-319)
            t2, t3 = t0 + dt, t0 + 0.5 * dt
            if self.precision != "fp32k":
                self._stage(q, r2, q, cfl=cs, stage_t=t2, Q=dq_dt, s=dt, y_mode=K.Y_AXPY,
                            out_mode=K.OUT_SSP33_2, A=q)
                self._stage(q, out_q[0], r2, cfl=cs, stage_t=t3, out_mode=K.OUT_SSP33_3, A=q)
            else:
                aux = self._buf("aux", q)
                self._stage(q, r2, q, cfl=cs, stage_t=t2, Q=dq_dt, s=dt, y_mode=K.Y_AXPY,
                            out_mode=K.OUT_F32_S2, A=q, B=dq_dt, aux_out=aux)
                self._stage(q, out_q[0], r2, cfl=cs, stage_t=t3, out_mode=K.OUT_F32_S3, A=q,
                            aux_in=aux, comp_in=comp, comp_out=out_q[1])
        else:   # SSP104
            # s1.t: t + dt/6., then += dt/6. per stage; t + dt/3. after the middle
            # update (sharpclaw/solver.py:409-427)
            s1a, s1b, qm = self._buf("s1a", q), self._buf("s1b", q), self._buf("qm", q)
            ts = t0 + dt / 6.
            self._stage(q, s1a, q, cfl=cs, stage_t=ts, Q=dq_dt, s=dt, y_mode=K.Y_AXPY6,
                        out_mode=K.OUT_ADD6)
            ts = ts + dt / 6.
            cur, nxt = s1a, s1b
            for _ in range(3):
                self._stage(q, nxt, cur, cfl=cs, stage_t=ts, out_mode=K.OUT_ADD6)
                ts = ts + dt / 6.
                cur, nxt = nxt, cur
            qm.copy_(q)
            K.ssp104_mid(qm, cur, fast=self._fast)
            ts = t0 + dt / 3.
            for _ in range(4):
                self._stage(q, nxt, cur, cfl=cs, stage_t=ts, out_mode=K.OUT_ADD6)
                ts = ts + dt / 6.
                cur, nxt = nxt, cur
            self._stage(q, out_q[0], cur, cfl=cs, stage_t=ts, out_mode=K.OUT_SSP104_FIN, A=qm)

    def _full_weno_step(self, q, comp, out_q, cs):
        """The same step (same dt) with full WENO5 at every interface: the first
        operator evaluation and the later stages, into out_q; Courant numbers
        max-reduced into cs."""
        flags, self._flags = self._flags, None
        try:
            dqdt = self._buf("dq_dt_full", q)
            self._stage(q, dqdt, q, cfl=cs, out_mode=K.OUT_DQDT)
            self._rk_stages(q, comp, dqdt, out_q, self.dt, cs)
        finally:
            self._flags = flags

    def _positive_fields(self):
        """Fields that must stay positive: depth (shallow water), density and
        energy (Euler)."""
        return {2: (0, 3), 3: (0,), 9: (0,)}.get(getattr(self.rp, "kind", None), ())

    def _judge(self, q, comp, out_q, sb):
        """Full WENO5 judges the indicator's step.  The step is recomputed with
        full WENO5 when it leaves a non-positive depth / density / energy (or a
        non-finite value), and every indicator.audit_every steps it is compared
        with the full-WENO5 step: above indicator.audit_tol (relative L1 of the
        states) the full-WENO5 result is kept and the flags are refreshed every
        step from then on.  Returns the buffers holding the accepted result."""
        ind = self.indicator
        ind.judge_stats["steps"] += 1
        pos = self._positive_fields()
        audit = ind.audit_every and ind.judge_stats["steps"] % int(ind.audit_every) == 0
        if pos:
            # only the fields up to the last checked one are read
            nf = max(pos) + 1
            K.state_sums(out_q[0][:nf], pos, out=sb[1:nf + 2])
            vals = self._read_back(sb)
            cfl_stages, nbad = vals[0], vals[nf + 1]
        else:
            cfl_stages, nbad = self._read_back(sb[0:1])[0], 0.0
        self.cfl.update_global_max(cfl_stages)
        if nbad > 0:
            ind.judge_stats["positivity_fallbacks"] += 1
            cs2 = self._stage_cfl()
            self._full_weno_step(q, comp, out_q, cs2)
            self.cfl.update_global_max(self._read_back(cs2)[0])
            return out_q
        if audit:
            alt = (self._buf("audit_q", q), self._buf("audit_comp", comp) if comp is not None
                   else None)
            cs2 = self._stage_cfl()
            self._full_weno_step(q, comp, alt, cs2)
            d = K.l1_diff(out_q[0], alt[0])
            (cfl_full,), (diff, norm) = self._read_back(cs2), self._read_back(d)
            rel = diff / norm if norm > 0 else diff
            ind.judge_stats["audits"] += 1
            ind.judge_stats["last_audit_rel_l1"] = rel
            ind.judge_stats["max_audit_rel_l1"] = max(ind.judge_stats["max_audit_rel_l1"], rel)
            if not rel <= ind.audit_tol:
                ind.judge_stats["audit_failures"] += 1
                ind.every = 1
                self.cfl.update_global_max(cfl_full)
                # the full-WENO5 result becomes the step's; its buffers swap
                # places with the scratch ones
                self._bufs["audit_q"] = out_q[0]
                if comp is not None:
                    self._bufs["audit_comp"] = out_q[1]
                return alt
        return out_q

    # ---- device-resident SSP33 loop
    def _device_loop_ok(self, solution):
        lo, hi = self._bc_codes()
        return (self.device_loop and self.time_integrator == "SSP33" and self.dt_variable
                and self.before_step is None and self.dq_src is None
                and not self.call_before_step_each_stage and BC.custom not in lo
                and BC.custom not in hi and getattr(self.rp, "num_aux", 0) == 0)

    def evolve_to_time(self, solution, tend=None):
        """Solver.evolve_to_time (solver.py:554-666); towards a given tend, SSP33
        without host callbacks runs on the device (same arithmetic, decisions and
        dt sequence)."""
        if not self._is_set_up:
            self.setup(solution)
        if tend is None or not self._device_loop_ok(solution):
            return super(SharpClawSolver, self).evolve_to_time(solution, tend)
        tstart = solution.t
        if self.cfl_desired > self.cfl_max:
            raise Exception("Variable time-stepping and desired CFL > maximum CFL")
        if tend <= tstart:
            self.logger.info("Already at or beyond end time: no evolution required.")
            return self.status
        _, accepted = self._run_device(solution, tstart, tend, self.max_steps)
        if solution.t < tend and accepted == self.max_steps:
            raise Exception("Maximum number of timesteps have been taken")
        return self.status

    def run_steps(self, solution, nsteps):
        """Attempt exactly nsteps SSP33 steps (no end time) in the device loop."""
        if not self._is_set_up:
            self.setup(solution)
        if not self._device_loop_ok(solution):
            raise ValueError("run_steps needs the device loop: SSP33, variable dt, no host "
                             "callbacks, built-in boundary conditions, no aux fields")
        return self._run_device(solution, solution.t, float("inf"), int(nsteps))

    def _run_device(self, solution, tstart, tend, max_attempts):
        """Up to max_attempts attempts towards tend; the indicator's audit
        attempts (every audit_every judged steps) run on the host path.  Returns
        (attempts, accepted)."""
        ind = self.indicator
        attempts = accepted = 0
        done = False
        while attempts < max_attempts and not done:
            n = max_attempts - attempts
            if ind is not None and ind.audit_every:
                every = int(ind.audit_every)
                k = every - ind.judge_stats["steps"] % every   # the k-th next one audits
                if k == 1:
                    acc, done = self._host_attempt(solution, tstart, tend)
                    attempts += 1
                    accepted += acc
                    continue
                n = min(n, k - 1)
            a, acc, done = self._device_attempts(solution, tstart, tend, n)
            attempts += a
            accepted += acc
        return attempts, accepted

    def time_stages(self, solution, nsteps):
        """nsteps attempts of the device loop launched without graphs, one at a
        time, with HIP events around the march launch of every main stage (on the
        stream it runs on).  Returns the march durations in ms of the stages that
        ran (a rejected attempt's successor skips stage 1; indicator audit
        attempts run on the host path and are not timed)."""
        n = int(nsteps)
        events = [torch.cuda.Event(enable_timing=True) for _ in range(6 * n)]
        st = torch.cuda.current_stream()
        for e in events:
            e.record(st)   # creates the event handles
        use, self.use_graph = self.use_graph, False
        self._stage_events, self._stage_ev_pos, self._stage_ran = events, 0, []
        try:
            self.run_steps(solution, n)
            used, ran = self._stage_ev_pos, self._stage_ran
        finally:
            self.use_graph = use
            self._stage_events, self._stage_ev_pos, self._stage_ran = None, 0, []
        torch.cuda.synchronize()
        return [events[2 * i].elapsed_time(events[2 * i + 1]) for i in range(used // 2)
                if ran[i]]

    def _host_attempt(self, solution, tstart, tend):
        """One attempt of Solver.evolve_to_time's loop (solver.py:607-651) on the
        host path; returns (accepted 0/1, reached tend)."""
        state = solution.state
        backup = self._backup(state)
        told = solution.t
        self.step(solution, False, tstart, tend)
        cfl = self.cfl.get_cached_max()
        self.accept_step = self.accept_reject_step(state)
        if self.accept_step:
            self.status["cflmax"] = max(cfl, self.status["cflmax"])
            solution.t += self.dt
            self.status["numsteps"] += 1
            return 1, solution.t >= tend
        self._restore(state, backup)
        solution.t = told
        return 0, False

    def _device_setup(self, solution):
        """The device loop's buffers and launch arguments for the state's current
        buffers: a ping-pong pair whose identities captured graphs keep, the
        work buffers, and on the indicator path its flags, their list and the
        full-WENO5 edge buffer (allocated when missing).  Returns
        (bufs, cur, args, kw, dq_dt buffer)."""
        state = solution.states[0]
        self._grid = state.grid
        self._state = state
        ind = self.indicator
        qin = state.device_q(self._dtype_code)
        other = self._take_scratch(qin)
        bufs, cur = (qin, other), 0
        for ent in self._graphs.values():
            b = ent["bufs"]
            if b[0][0] is qin[0] or b[1][0] is qin[0]:
                bufs = b
                cur = 0 if b[0][0] is qin[0] else 1
                break
        q = bufs[0][0]
        fp32 = self.precision == "fp32k"
        dqdt, r2 = self._buf("ctl_dqdt", q), self._buf("ctl_r2", q)
        aux23 = self._buf("ctl_aux23", q) if fp32 else None
        kw = {}
        if ind is not None:
            meqn, my, mx = q.shape
            nint = K.mlp_num_interfaces(mx, my)
            if ind._flags is None or ind._flags.numel() != nint or ind._list is None:
                ind._flags = torch.empty(nint, dtype=torch.uint8, device=q.device)
                ind._list = (torch.empty(nint, dtype=torch.int32, device=q.device),
                             torch.zeros(1, dtype=torch.int32, device=q.device))
                ind._age = None   # nothing computed yet: the next attempt refreshes
            kw = dict(indicator=ind, dq_dt_full=self._buf("ctl_dqdt_fb", q),
                      recon=ind.recon_buffer(q), positive=self._positive_fields())
        if self._ctl is None:
            self._ctl = torch.zeros(ctypes_sizeof_sharpctl(), dtype=torch.uint8, device="cuda")
        lo, hi = self._bc_codes()
        dx, dy = self._grid.delta
        comps = (bufs[0][1], bufs[1][1])
        args = (self._rp_struct, self.char_decomp, (bufs[0][0], bufs[1][0]), comps, self._ctl,
                lo, hi, self.reflect_index, dx, dy, dqdt, r2, aux23)
        kw["fast"] = self._fast
        return bufs, cur, args, kw, dqdt

    def capture_graphs(self, solution, nsteps):
        """Capture (without running) the graphs run_steps(solution, nsteps) will
        replay, so that a timed run holds no capture."""
        if not self._is_set_up:
            self.setup(solution)
        if not (self.use_graph and self._device_loop_ok(solution)):
            return
        bufs, cur, args, kw, _ = self._device_setup(solution)
        for m in sorted(self._chunk_sizes(int(nsteps))):
            self._graph(bufs, args, kw, m)
        self._scratch = bufs[1 - cur]

    def _chunk_sizes(self, nsteps):
        """The launch sizes run_steps(solution, nsteps) uses: device runs up to
        each indicator audit (an audit attempt runs on the host path), each cut
        into chunks of at most graph_chunk attempts."""
        ind = self.indicator
        every = int(ind.audit_every) if ind is not None and ind.audit_every else 0
        steps = ind.judge_stats["steps"] if ind is not None else 0
        sizes, attempts = set(), 0
        while attempts < nsteps:
            n = nsteps - attempts
            if every:
                k = every - steps % every
                if k == 1:
                    attempts += 1
                    steps += 1
                    continue
                n = min(n, k - 1)
            left = n
            while left > 0:
                m = min(self.graph_chunk, left)
                sizes.add(m)
                left -= m
            attempts += n
            steps += n
        return sizes

    def _device_attempts(self, solution, tstart, tend, n):
        """n attempts in the device loop; the solver, the state and the
        indicator take the controller's state afterwards."""
        import ctypes
        from . import _lib as L
        state = solution.states[0]
        ind = self.indicator
        bufs, cur, args, kw, dqdt = self._device_setup(solution)
        if not self.accept_step:
            dqdt.copy_(self.dq_dt)   # the rate a rejected attempt's successor reuses
            # that successor skips stage 1: its get_dt and the CFL rescale
            # (sharpclaw/solver.py:296-303) happen here
            self.get_dt(solution.t, tstart, tend, False)
            cfl = self.cfl.get_cached_max()
            self.cfl.set_global_max(self.dt / self.dt_old * cfl)
            self.dt_old = self.dt
        c = L.SharpCtl()
        c.t, c.dt, c.tstart, c.tend = solution.t, self.dt, tstart, tend
        c.have_dt_old = int(self.dt_old is not None)
        c.dt_old = self.dt_old if self.dt_old is not None else 0.0
        cached = self.cfl.get_cached_max()
        c.cfl = 0.0 if cached is None else cached   # None before the first step: stage 1 sets it
        c.cfl_desired, c.cfl_max, c.dt_max = self.cfl_desired, self.cfl_max, self.dt_max
        c.cflmax, c.dtmin, c.dtmax = (self.status["cflmax"], self.status["dtmin"],
                                      self.status["dtmax"])
        c.accept_prev = int(bool(self.accept_step))
        c.cur = cur
        c.max_steps = int(n)
        c.run1, c.run23 = c.accept_prev, 1
        if ind is not None:
            # the first attempt's flag refresh (step_flags); later ones on the device
            every = ind.refresh_interval(self.cfl_max)
            age = ind._age
            refresh = age is None or age >= every
            if refresh:
                age = 0
            c.age, c.every, c.run_flags = age + 1, every, int(refresh)
            c.judge = 1
            c.judge_steps = ind.judge_stats["steps"]
            c.fallbacks = ind.judge_stats["positivity_fallbacks"]
        self._ctl.copy_(torch.frombuffer(bytearray(bytes(c)), dtype=torch.uint8))
        host = L.SharpCtl.from_buffer_copy(bytes(c))
        timing = self._stage_events is not None
        while not host.done and host.attempts < n:
            # as many attempts as the remaining time needs (estimated from the
            # current dt) before reading the controller back; a launch whose
            # judge falls back to full WENO5 ends early, so the controller's
            # attempt count, not the launched one, says what is left.  Timed
            # runs launch one attempt at a time and keep the gates it ran with.
            before = host.attempts
            todo = _attempts_to_launch(n - host.attempts, tend, host.t, host.dt,
                                       self.graph_chunk)
            if timing:
                todo = 1
                self._stage_ran.extend([bool(host.run1), bool(host.run23), bool(host.run23)])
            while todo > 0:
                m = min(self.graph_chunk, todo)
                if self.use_graph:
                    self._graph(bufs, args, kw, m).replay()
                elif timing:
                    p0 = self._stage_ev_pos
                    K.sharp2_ssp33_run(*args, m, stage_events=self._stage_events[p0:p0 + 6 * m],
                                       **kw)
                    self._stage_ev_pos = p0 + 6 * m
                else:
                    K.sharp2_ssp33_run(*args, m, **kw)
                todo -= m
            ctypes.memmove(ctypes.addressof(host), bytes(self._ctl.cpu().numpy().tobytes()),
                           ctypes.sizeof(host))
            if not host.done and host.attempts == before:
                raise RuntimeError("device SSP33 loop made no progress (controller attempts "
                                   f"stayed at {before})")
        solution.t = host.t
        self.dt = host.dt
        self.dt_old = host.dt_old if host.have_dt_old else None
        self.cfl.set_global_max(host.cfl)
        self.status["cflmax"], self.status["dtmin"], self.status["dtmax"] = (host.cflmax,
                                                                              host.dtmin, host.dtmax)
        self.status["numsteps"] += host.accepted
        self.accept_step = bool(host.accept_prev)
        self.dq_dt = dqdt
        state.set_device_q(*bufs[host.cur])
        self._scratch = bufs[1 - host.cur]
        if ind is not None:
            ind._age = host.age
            ind.judge_stats["steps"] = host.judge_steps
            ind.judge_stats["positivity_fallbacks"] = host.fallbacks
        return host.attempts, host.accepted, host.done == 1

    @staticmethod
    def _graph_key(args, kw, n):
        """What a captured run depends on: the address of every buffer it
        touches, the indicator's scalars and the attempt count."""
        key = [n]
        for x in list(args) + [kw.get(k) for k in sorted(kw)]:
            if isinstance(x, torch.Tensor):
                key.append(x.data_ptr())
            elif isinstance(x, (tuple, list)):
                key.extend(t.data_ptr() if isinstance(t, torch.Tensor) else t for t in x)
            elif x is not None and hasattr(x, "params"):   # the indicator
                key.extend([x.params.data_ptr(), x._flags.data_ptr(), x._list[0].data_ptr(),
                            x._list[1].data_ptr(), x.flat, x.threshold])
            elif isinstance(x, (int, float, str)) or x is None:
                key.append(x)
        return tuple(key)

    def _graph(self, bufs, args, kw, n):
        """The captured graph of n attempts with these arguments (captured on
        first use)."""
        key = self._graph_key(args, kw, n)
        ent = self._graphs.get(key)
        if ent is None:
            from .device import no_gc
            g = torch.cuda.CUDAGraph()
            st = torch.cuda.Stream()
            st.wait_stream(torch.cuda.current_stream())
            # no garbage collection while capturing: collecting an old graph
            # would destroy it mid-capture, which HIP refuses
            with no_gc(), torch.cuda.stream(st):
                with torch.cuda.graph(g, stream=st):
                    K.sharp2_ssp33_run(*args, n, **kw)
            torch.cuda.current_stream().wait_stream(st)
            ent = {"graph": g, "bufs": bufs, "args": args, "kw": kw}
            self._graphs[key] = ent
        return ent["graph"]

    def _lmm(self, *args, **kw):
        """K.lmm_update in this solver's arithmetic variant."""
        K.lmm_update(*args, fast=self._fast, **kw)

    def dq(self, state):
        """dt*L(q) on the interior as a host array (sharpclaw/solver.py:532-542)."""
        from .device import q_from_device
        self._grid = state.grid
        self._state = state
        q, _ = state.device_q(self._dtype_code)
        self._qaux = self._aux_input(state) if self._reads_aux(state) else None
        out = torch.empty_like(q)
        self._stage(q, out, q, out_mode=K.OUT_D)
        return q_from_device(out)


class SharpClawSolver2D(SharpClawSolver):
    """sharpclaw/solver.py:829-898"""

    def __init__(self, riemann_solver=None, claw_package=None):
        self.num_dim = 2
        self.reflect_index = [1, 2]
        self._grid = None
        super(SharpClawSolver2D, self).__init__(riemann_solver, claw_package)
