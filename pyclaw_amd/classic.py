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
        self._method[6] = state.num_aux

    def setup(self, solution):
        self._check_cfl_settings()
        self._set_mthlim()
        if self.fwave:
            raise NotImplementedError("f-wave Riemann solvers are not available on the device")
        state = solution.states[0]
        if state.num_aux > 0 or state.index_capa >= 0:
            raise NotImplementedError("aux arrays and capacity functions are not supported on the device")
        self._set_method(state)
        if self.rp is None or not hasattr(self.rp, "kind"):
            raise Exception("the solver needs one of the pyclaw_amd.riemann modules")
        self._rp_struct = K.make_rp(self.rp.kind, state.problem_data)
        self._is_set_up = True


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
        if BC.custom in lo or BC.custom in hi:
            src, gin = self._custom_bc_input(state, lo, hi)
            lo, hi = [4, 4], [4, 4]
        if self.dimensional_split:
            self._work = K.step2_split(self._rp_struct, src[0], out[0], cfl_dev, lo, hi,
                                       self.reflect_index, self._method, self._mthlim, dx, dy,
                                       self.dt, comp_in=src[1], comp_out=out[1], gin=gin,
                                       num_ghost=self.num_ghost, work=self._work)
        else:
            K.step2_unsplit(self._rp_struct, src[0], out[0], cfl_dev, lo, hi, self.reflect_index,
                            self._method, self._mthlim, dx, dy, self.dt, comp_in=src[1],
                            comp_out=out[1], gin=gin, num_ghost=self.num_ghost)
        cfl = float(cfl_dev.item())
        self.cfl.update_global_max(cfl)
        state.set_device_q(*out)
        self._scratch = qin

    def _custom_bc_input(self, state, lo, hi):
        """User BC functions act on a host qbc (solver.py:393-398); the kernel then
        reads the stored ghosts (BC_GIVEN)."""
        import torch
        from . import _lib as L
        from .device import to_planar, torch_dtype
        ng = self.num_ghost
        q = state.q
        qbc = np.zeros((q.shape[0],) + tuple(n + 2 * ng for n in q.shape[1:]), order="F")
        qbc[(slice(None),) + (slice(ng, -ng),) * self.num_dim] = q
        _fill_bcs_host(self, state, qbc, lo, hi)
        t = torch.from_numpy(to_planar(qbc)).to("cuda", torch_dtype(self._dtype_code))
        comp = state.device_q(self._dtype_code)[1]
        if comp is not None:
            cp = torch.zeros_like(t)
            cp[:, ng:-ng, ng:-ng] = comp
            comp = cp
        return (t, comp), L.geom2(q.shape[1], q.shape[2], ng)


class ClawSolver3D(ClawSolver):
    """classic/solver.py:554-706 (dimension-split path)."""
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
            raise NotImplementedError("the unsplit 3D step (step3.f90) is not part of pyclaw_amd")
        super(ClawSolver3D, self).setup(solution)

    def step_hyperbolic(self, solution):
        state = solution.states[0]
        dx, dy, dz = state.grid.delta
        lo, hi = self._bc_codes()
        if BC.custom in lo or BC.custom in hi:
            raise NotImplementedError("custom boundary conditions in 3D are not supported")
        qin = state.device_q(self._dtype_code)
        out = self._take_scratch(qin)
        cfl_dev = self._cfl_buffer()
        K.cfl_reset(cfl_dev)
        self._buf3 = K.step3_split(self._rp_struct, qin[0], out[0], cfl_dev, lo, hi,
                                   self.reflect_index, self._method, self._mthlim, dx, dy, dz,
                                   self.dt, comp_in=qin[1], comp_out=out[1], work=self._buf3,
                                   num_ghost=self.num_ghost)
        self.cfl.update_global_max(float(cfl_dev.item()))
        state.set_device_q(*out)
        self._scratch = qin


def _fill_bcs_host(solver, state, qbc, lo, hi):
    """Solver._apply_bcs on a host qbc (solver.py:351-519), for custom BCs."""
    ng = solver.num_ghost
    for idim, dim in enumerate(state.grid.dimensions):
        for upper, codes in ((False, lo), (True, hi)):
            b = codes[idim]
            if b == BC.custom:
                fun = solver.user_bc_upper if upper else solver.user_bc_lower
                fun(state, dim, state.t, qbc, solver.auxbc, ng)
                continue
            arr = np.moveaxis(qbc, idim + 1, 1)
            r = solver.reflect_index[idim]
            for i in range(ng):
                if not upper:
                    if b == BC.extrap:
                        arr[:, i, ...] = arr[:, ng, ...]
                    elif b == BC.periodic:
                        arr[:, i, ...] = arr[:, -2 * ng + i, ...]
                    elif b == BC.wall:
                        arr[:, i, ...] = arr[:, 2 * ng - 1 - i, ...]
                        arr[r, i, ...] = -arr[r, 2 * ng - 1 - i, ...]
                else:
                    if b == BC.extrap:
                        arr[:, -i - 1, ...] = arr[:, -ng - 1, ...]
                    elif b == BC.periodic:
                        arr[:, -ng + i, ...] = arr[:, ng + i, ...]
                    elif b == BC.wall:
                        arr[:, -i - 1, ...] = arr[:, -2 * ng + i, ...]
                        arr[r, -i - 1, ...] = -arr[r, -2 * ng + i, ...]
