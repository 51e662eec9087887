"""Oracle time loop: a restatement of PyClaw's Python-side stepping semantics
driving the oracle's FP64 kernels.

TEST INFRASTRUCTURE ONLY (never imported by ``pyclaw_amd``).  It follows:
  Controller.run output times          controller.py:275-395 (output_style 1)
  Solver.evolve_to_time                solver.py:554-666
  Solver.get_dt / get_dt_new           solver.py:531-552
  ClawSolver.step                      classic/solver.py:96-146
  ClawSolver2D/3D.step_hyperbolic      classic/solver.py:498-549, 648-706
  SharpClawSolver.step (Euler, SSP22, SSP33, SSP104, SSPLMMk2/k3)
                                       sharpclaw/solver.py:283-430
  SharpClawSolver.update_saved_values  sharpclaw/solver.py:432-490
  SharpClawSolver.accept_reject_step   sharpclaw/solver.py:622-650
  SharpClawSolver.get_dt_new           sharpclaw/solver.py:653-685
  SharpClawSolver2D.dq_hyperbolic      sharpclaw/solver.py:848-898
"""
import numpy as np

from . import oracle as O


class OracleRun:
    """One problem: interior q (meqn, mx, my[, mz]) plus method settings."""

    def __init__(self, rp, q, delta, bc_lower, bc_upper, solver="classic",
                 dimensional_split=True, transverse=1, limiters=1, order=2,
                 cfl_desired=0.9, cfl_max=1.0, dt_initial=0.1, dt_max=1e99,
                 max_steps=10000, dt_variable=True, time_integrator="SSP104",
                 char_decomp=0, num_ghost=None, aux=None, aux_bc_lower=None,
                 aux_bc_upper=None, before_step=None, step_source=None, source_split=1,
                 index_capa=-1, auxbc_fn=None, lmm_steps=4, user_bc_lower=None,
                 rk_a=None, rk_b=None, lmm_alpha=None, lmm_beta=None, dq_src=None,
                 rk_c=None, fwave=False):
        self.rp = rp
        self.ndim = q.ndim - 1
        self.delta = [float(d) for d in delta]
        self.bc_lower, self.bc_upper = list(bc_lower), list(bc_upper)
        self.solver = solver
        self.split = bool(dimensional_split)
        self.transverse = transverse
        self.order = order
        self.cfl_desired, self.cfl_max = cfl_desired, cfl_max
        self.dt_initial, self.dt_max = dt_initial, dt_max
        self.max_steps, self.dt_variable = max_steps, dt_variable
        self.time_integrator = time_integrator
        self.char_decomp = char_decomp
        mw = O.MWAVES[rp.kind]
        lim = limiters if isinstance(limiters, list) else [limiters]
        self.mthlim = lim * mw if len(lim) == 1 else lim
        if num_ghost is None:
            num_ghost = 2 if solver == "classic" else 3
        self.ng = num_ghost
        self.q = np.asfortranarray(q, dtype=np.float64).copy(order="F")
        self.qbc = np.zeros((q.shape[0],) + tuple(n + 2 * self.ng for n in q.shape[1:]), order="F")
        self.t = 0.0
        self.dt = dt_initial
        self.dt_old = None
        self.cfl = 0.0              # the CFL object's cached max
        self.accept_step = True
        self.dq_dt = None
        self.numsteps = 0
        self.dt_history = []
        # aux fields (interior, (maux, mx, my)), their BCs (no negation), and the
        # user callbacks of Solver.evolve_to_time / ClawSolver.step
        self.aux = None if aux is None else np.asfortranarray(aux, dtype=np.float64).copy(order="F")
        self.aux_bc_lower, self.aux_bc_upper = aux_bc_lower, aux_bc_upper
        self.before_step, self.step_source, self.source_split = before_step, step_source, source_split
        # capacity field (0-based aux index, -1: none); auxbc_fn(run) -> the aux
        # array with its ghosts filled, for custom aux boundary conditions
        self.index_capa, self.auxbc_fn = index_capa, auxbc_fn
        # user_bc_lower(run, dim, qbc): Solver.user_bc_lower for BC.custom lower
        # sides, called in the reference's per-dimension order
        self.user_bc_lower = user_bc_lower
        # dq_src(run, q, dt) -> the source increment SharpClaw adds to dq
        self.dq_src = dq_src
        # the Riemann solver returns f-waves (ClawSolver.fwave -> step2's use_fwave)
        self.fwave = bool(fwave)
        # multistep state (SSPLMMk2 / k3)
        self.lmm_steps = lmm_steps
        self.sspcoeff = None
        if time_integrator == "SSPLMMk2":
            self.sspcoeff = (lmm_steps - 2.) / (lmm_steps - 1.)
        elif time_integrator == "SSPLMMk3":
            self.sspcoeff = (lmm_steps - 3.) / (lmm_steps - 1.)
        self.sspcoeff0 = None
        # the general methods: RK (Butcher a, b) and LMM (alpha, beta)
        self.rk_a, self.rk_b, self.rk_c = rk_a, rk_b, rk_c
        self.stage_t = 0.0
        self.plant_stage_time = False
        self.plant_lmm3 = False
        self.lmm_alpha, self.lmm_beta = lmm_alpha, lmm_beta
        # number of saved states (sharpclaw/solver.py:596-610)
        self.nreg = len(lmm_alpha) if time_integrator == "LMM" else lmm_steps
        self.registers = [None] * self.nreg
        self.prev_dq_dt, self.prev_dt, self.prev_dtFE = [], [], []

    # ---- boundary conditions: Solver._apply_bcs (solver.py:351-437)
    def _apply_bcs(self, q):
        ng = self.ng
        sl = (slice(None),) + tuple(slice(ng, -ng) for _ in range(self.ndim))
        self.qbc[sl] = q
        if self.user_bc_lower is not None:
            # dimension by dimension, lower side then upper (solver.py:351-367)
            nd = self.ndim
            for d in range(nd):
                lo, hi = [O.CUSTOM] * nd, [O.CUSTOM] * nd
                lo[d], hi[d] = self.bc_lower[d], self.bc_upper[d]
                if lo[d] == O.CUSTOM:
                    self.user_bc_lower(self, d, self.qbc)
                (O.bc2 if nd == 2 else O.bc3)(self.qbc, ng, lo, hi)
        elif self.ndim == 2:
            O.bc2(self.qbc, ng, self.bc_lower, self.bc_upper)
        else:
            O.bc3(self.qbc, ng, self.bc_lower, self.bc_upper)
        return self.qbc

    def _interior(self, a):
        ng = self.ng
        sl = (slice(None),) + tuple(slice(ng, -ng) for _ in range(self.ndim))
        return np.asfortranarray(a[sl]).copy(order="F")

    # ---- Solver.get_dt (solver.py:535-552)
    def _get_dt(self, t, tstart, tend, take_one_step):
        cfl = self.cfl
        if self.dt_variable and self.dt_old is not None:
            if cfl > 0.0:
                if self.solver == "classic":
                    self.dt = min(self.dt_max, self.dt * self.cfl_desired / cfl)
                else:
                    self._get_dt_new_sharp(cfl)
            else:
                self.dt = self.dt_max
        else:
            self.dt_old = self.dt
        if not take_one_step:
            if t + self.dt > tend and tstart < tend:
                self.dt = tend - t
            if tend - t - self.dt < 1.e-14 * t:
                self.dt = tend - t

    # ---- SharpClawSolver.get_dt_new (sharpclaw/solver.py:653-685)
    def _get_dt_new_sharp(self, cfl):
        self.dt_old = self.dt
        if "SSPLMM" in self.time_integrator:
            step_index = self.numsteps + 1
            if step_index < self.lmm_steps:
                sspcoeff_ratio = self.sspcoeff0 / self.sspcoeff
                self.dt = sspcoeff_ratio * self.dt * self.cfl_desired / cfl
            elif self.accept_step:
                s = self.lmm_steps
                p = int(self.time_integrator[-1])
                mu = min([self.prev_dtFE[i] for i in range(s)])
                H = sum(self.prev_dt[1:])
                self.dt = H * mu / (H + (p - 1) * mu)
        else:
            self.dt = self.dt * self.cfl_desired / cfl

    # ---- SharpClawSolver.accept_reject_step (sharpclaw/solver.py:622-650)
    def _accept(self):
        if self.solver == "sharp" and "LMM" in self.time_integrator:
            if self.numsteps + 1 < self.nreg:
                ratio = 1. if self.time_integrator == "LMM" else self.sspcoeff0 / self.sspcoeff
                return not (self.cfl > ratio * self.cfl_max)
            return True
        return not (self.cfl > self.cfl_max)

    # ---- SharpClawSolver.update_saved_values (sharpclaw/solver.py:432-490)
    def _update_saved_values(self, step_index):
        if self.prev_dt == []:
            self.registers[-1] = self.q.copy(order="F")
            self.prev_dq_dt.append(self.dq_dt)
            self.prev_dt.append(0.)
            if "SSPLMM" in self.time_integrator:
                dtFE = self.dt / self.cfl * self.cfl_max / self.sspcoeff
                self.prev_dtFE.append(dtFE)
                self.sspcoeff0 = 1.0   # _sspcoeff['SSP22']
        elif self.accept_step and self.time_integrator == "LMM":
            if step_index <= self.nreg:
                self.prev_dq_dt.append(self.dq_dt)
            else:
                self.prev_dq_dt = self.prev_dq_dt[1:] + self.prev_dq_dt[:1]
                self.prev_dq_dt[-1] = self.dq_dt
            self.registers = self.registers[1:] + self.registers[:1]
            self.registers[-1] = self.q.copy(order="F")
        elif self.accept_step:
            dtFE = self.dt / self.cfl * self.cfl_max / self.sspcoeff
            if step_index <= self.lmm_steps:
                if self.time_integrator == "SSPLMMk3":
                    self.prev_dq_dt.append(self.dq_dt)
                self.prev_dt.append(self.dt_old)
                self.prev_dtFE.append(dtFE)
            else:
                if self.time_integrator == "SSPLMMk3":
                    self.prev_dq_dt = self.prev_dq_dt[1:] + self.prev_dq_dt[:1]
                    self.prev_dq_dt[-1] = self.dq_dt
                self.prev_dt = self.prev_dt[1:] + self.prev_dt[:1]
                self.prev_dt[-1] = self.dt_old
                self.prev_dtFE = self.prev_dtFE[1:] + self.prev_dtFE[:1]
                self.prev_dtFE[-1] = dtFE
            self.registers = self.registers[1:] + self.registers[:1]
            self.registers[-1] = self.q.copy(order="F")
        return step_index

    def _auxbc(self):
        if self.auxbc_fn is not None:
            return np.asfortranarray(self.auxbc_fn(self), dtype=np.float64)
        ng = self.ng
        abc = np.zeros((self.aux.shape[0],) + tuple(n + 2 * ng for n in self.aux.shape[1:]),
                       order="F")
        abc[(slice(None),) + (slice(ng, -ng),) * 2] = self.aux
        O.bc2(abc, ng, self.aux_bc_lower, self.aux_bc_upper, reflect=(-1, -1))
        return abc

    # ---- classic step (classic/solver.py:96-146, 498-549, 648-706)
    def _step_classic(self, t, tstart, tend, take_one_step):
        self._get_dt(t, tstart, tend, take_one_step)
        self.cfl = 0.0
        if self.source_split == 2 and self.step_source is not None:
            self.step_source(self, self.dt / 2.0)
        qbc = self._apply_bcs(self.q)
        qold = qbc.copy(order="F")
        ng, dt = self.ng, self.dt
        if self.ndim == 2:
            dx, dy = self.delta
            if self.split:
                maux = 0 if self.aux is None else self.aux.shape[0]
                abc = None if self.aux is None else self._auxbc()
                meth = O.method_array(order=self.order, dimensional_split=True,
                                      dt_variable=self.dt_variable, num_aux=maux)
                c1 = O.step2(self.rp, qold, qbc, ng, dx, dy, dt, meth, self.mthlim, ids=1,
                             use_fwave=self.fwave, aux=abc)
                c2 = O.step2(self.rp, qbc, qbc, ng, dx, dy, dt, meth, self.mthlim, ids=2,
                             use_fwave=self.fwave, aux=abc)
                cfl = max(c1, c2)
            else:
                maux = 0 if self.aux is None else self.aux.shape[0]
                abc = None if self.aux is None else self._auxbc()
                meth = O.method_array(order=self.order, transverse=self.transverse,
                                      dt_variable=self.dt_variable, num_aux=maux,
                                      index_capa=self.index_capa)
                cfl = O.step2(self.rp, qold, qbc, ng, dx, dy, dt, meth, self.mthlim, ids=0,
                              use_fwave=self.fwave, aux=abc)
        else:
            dx, dy, dz = self.delta
            if self.split:
                meth = O.method_array(order=self.order, dimensional_split=True,
                                      dt_variable=self.dt_variable)
                c1 = O.step3ds(self.rp, qold, qbc, ng, dx, dy, dz, dt, meth, self.mthlim, 1)
                c2 = O.step3ds(self.rp, qbc, qbc, ng, dx, dy, dz, dt, meth, self.mthlim, 2)
                c3 = O.step3ds(self.rp, qbc, qbc, ng, dx, dy, dz, dt, meth, self.mthlim, 3)
                cfl = max(c1, c2, c3)
            else:   # step3 (classic/solver.py:694-699); transverse = method(3)
                meth = O.method_array(order=self.order, transverse=self.transverse,
                                      dt_variable=self.dt_variable)
                cfl = O.step3(self.rp, qold, qbc, ng, dx, dy, dz, dt, meth, self.mthlim)
        self.cfl = max(self.cfl, cfl)
        self.q = self._interior(qbc)
        if self.cfl >= self.cfl_max:
            return
        if self.step_source is not None:
            self.step_source(self, self.dt / 2.0 if self.source_split == 2 else self.dt)

    # ---- SharpClaw dq (sharpclaw/solver.py:848-898), returns dt*L(q) interior.
    # t: the time of the state dq acts on (the register's t,
    # sharpclaw/solver.py:313-338, 409-427); callbacks read it as run.stage_t.
    def _dq(self, q, t=None):
        # plant_stage_time (tests only): every stage sees the step's start time
        self.stage_t = self.t if (t is None or self.plant_stage_time) else t
        qbc = self._apply_bcs(q)
        dx, dy = self.delta
        reads_aux = self.aux is not None and self.rp.kind in (6, 7, 8)
        dq, cfl = O.sharp_flux2(self.rp, qbc, self.ng, dx, dy, self.dt, self.char_decomp,
                                aux=self._auxbc() if reads_aux else None)
        self.cfl = max(self.cfl, cfl)
        dq = self._interior(dq)
        if self.dq_src is not None:
            # dq_hyperbolic + dq_src (sharpclaw/solver.py:537-540)
            dq = dq + self.dq_src(self, q, self.dt)
        return dq

    # ---- SharpClawSolver.step (sharpclaw/solver.py:283-430)
    def _step_sharp(self, t, tstart, tend, take_one_step):
        step_index = self.numsteps + 1
        if self.accept_step:
            self.cfl = 0.0
            self.dq_dt = self._dq(self.q) / self.dt
        if "LMM" in self.time_integrator:
            step_index = self._update_saved_values(step_index)
        self._get_dt(t, tstart, tend, take_one_step)
        cfl = self.cfl
        self.cfl = self.dt / self.dt_old * cfl
        self.dt_old = self.dt
        q, dt = self.q, self.dt
        t0 = self.t
        if self.time_integrator == "Euler":
            self.q = q + dt * self.dq_dt
        elif self.time_integrator == "SSP22" or ("SSPLMM" in self.time_integrator
                                                   and step_index < self.lmm_steps):
            r = q + dt * self.dq_dt
            self.q = 0.5 * (q + r + self._dq(r, t0 + dt))
        elif self.time_integrator == "SSPLMMk2":
            om = sum(self.prev_dt[1:]) / dt
            r = (om - 1.) / om
            delta = 1. / om ** 2
            beta = (om + 1.) / om
            self.q = beta * (r * q + dt * self.dq_dt) + delta * self.registers[-self.lmm_steps]
        elif self.time_integrator == "SSPLMMk3":
            om = sum(self.prev_dt[1:]) / dt
            omk = om + 1.
            r = (om - 2.) / om
            delta0 = (4 * omk - om ** 2) / om ** 3
            beta0 = omk / om ** 2
            bk = omk ** 2 / om ** 2
            # plant_lmm3 (tests only): the oldest rate replaced by the newest one
            old = self.prev_dq_dt[-1 if self.plant_lmm3 else -self.lmm_steps]
            self.q = (bk * (r * q + dt * self.prev_dq_dt[-1]) +
                      (r * beta0 + delta0) * self.registers[-self.lmm_steps] +
                      beta0 * dt * old)
        elif self.time_integrator == "SSP33":
            r = q + dt * self.dq_dt
            r = 0.75 * q + 0.25 * (r + self._dq(r, t0 + dt))
            self.q = 1. / 3. * q + 2. / 3. * (r + self._dq(r, t0 + 0.5 * dt))
        elif self.time_integrator == "RK":
            a, b = self.rk_a, self.rk_b
            regs = [dt * self.dq_dt]
            for i in range(1, len(b)):
                r = q.copy(order="F")
                for j in range(i):
                    r = r + a[i][j] * regs[j]
                regs.append(self._dq(r, t0 + dt * self.rk_c[i]))
            for j in range(len(b)):
                q = q + b[j] * regs[j]
            self.q = q
        elif self.time_integrator == "LMM" and step_index >= self.nreg:
            al, be, R, D = self.lmm_alpha, self.lmm_beta, self.registers, self.prev_dq_dt
            qn = al[-1] * R[-1] + be[-1] * dt * D[-1]
            for i in range(self.lmm_steps - 1):
                qn = qn + (al[i] * R[i] + be[i] * dt * D[i])
            self.q = qn
        elif self.time_integrator in ("SSP104", "LMM"):
            s1 = q + dt * self.dq_dt / 6.
            ts = t0 + dt / 6.
            for _ in range(4):
                s1 = s1 + self._dq(s1, ts) / 6.
                ts = ts + dt / 6.
            q = q / 25. + 0.36 * s1
            s1 = 15. * q - 5. * s1
            ts = t0 + dt / 3.
            for _ in range(4):
                s1 = s1 + self._dq(s1, ts) / 6.
                ts = ts + dt / 6.
            self.q = q + (0.6 * s1 + 0.1 * self._dq(s1, ts))
        else:
            raise ValueError(self.time_integrator)
        self.q = np.asfortranarray(self.q)

    # ---- Solver.evolve_to_time (solver.py:554-666)
    def evolve_to_time(self, tend):
        tstart = self.t
        take_one_step = tend is None
        max_steps = self.max_steps
        if not self.dt_variable:
            max_steps = 1 if take_one_step else int((tend - tstart + 1e-10) / self.dt)
        if self.dt_variable and self.cfl_desired > self.cfl_max:
            raise Exception("Variable time-stepping and desired CFL > maximum CFL")
        num_steps = 0
        for n in range(max_steps):
            q_backup = self.q.copy(order="F")
            told = self.t
            if self.before_step is not None:
                self.before_step(self)
            if self.solver == "classic":
                self._step_classic(self.t, tstart, tend, take_one_step)
            else:
                self._step_sharp(self.t, tstart, tend, take_one_step)
            self.accept_step = self._accept()
            if self.accept_step:
                if self.dt_variable:
                    self.t += self.dt
                else:
                    self.t = tstart + (n + 1) * self.dt
                num_steps += 1
                self.numsteps += 1
                self.dt_history.append(self.dt)
            else:
                if self.dt_variable:
                    self.q = q_backup
                    self.t = told
                else:
                    raise Exception("CFL too large, giving up!")
            if take_one_step or self.t >= tend:
                break
        if not take_one_step and self.dt_variable and self.t < tend and num_steps == max_steps:
            raise Exception("Maximum number of timesteps have been taken")

    # ---- Controller.run (controller.py:275-395), output_style 1
    def run(self, tfinal, num_output_times=10):
        self.dt = self.dt_initial
        frames = [self.q.copy(order="F")]
        for t in np.linspace(self.t, tfinal, num_output_times + 1)[1:]:
            self.evolve_to_time(t)
            frames.append(self.q.copy(order="F"))
        return frames


def cell_centers(lower, upper, n):
    """Dimension.centers (geometry.py:546-553): lower + (i+0.5)*delta."""
    delta = (upper - lower) / float(n)
    return np.array([lower + (i + 0.5) * delta for i in range(n)]), delta
