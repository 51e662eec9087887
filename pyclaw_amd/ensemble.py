"""Ensembles of independent 2D unsplit problems on one GPU.

Many ClawSolver2D problems of uneven grid sizes that share the Riemann solver,
method and boundary conditions are advanced by the same two launches per step
(clawhip_ensemble_run): each member keeps its own dt, CFL and accept/reject
decision in its own device step controller, with exactly the arithmetic of
ClawSolver2D.run_steps on that member alone.  Runs of steps are captured in a
HIP graph; the slices-per-wavefront shape is chosen by a timing autotuner.
There is no counterpart in the reference, which steps one grid per Solver.
"""
import ctypes
import time

import numpy as np
import torch

from . import _lib as L
from . import kernels as K
from .classic import ctypes_memmove
from .device import q_from_device, q_to_device


class Ensemble(object):
    """solver: a configured ClawSolver2D (unsplit; its Riemann solver, method,
    limiters, CFL settings, boundary conditions and precision apply to every
    member).  solutions: the members' Solutions (any grid sizes); each member
    starts from its Solution's q and t with dt = dt_initial (or `dts[k]`)."""

    def __init__(self, solver, solutions, dts=None):
        if solver.dimensional_split:
            raise ValueError("ensembles run the unsplit step")
        self.solver = solver
        self.solutions = list(solutions)
        self.dts = dts
        self.rows = 0
        self._table = None
        self._graphs = {}
        self._setup()

    # ---------------------------------------------------------------- set-up
    def _setup(self):
        s = self.solver
        first = self.solutions[0]
        if not s._is_set_up:
            s.setup(first)
        s._set_method(first.state)
        self._code = s._dtype_code
        # the kernels' arithmetic: fp64fast runs the FMA-contracted FP64 variant
        self._run_code = L.F64_FAST if s._fast else self._code
        lo, hi = s._bc_codes()
        if 0 in lo or 0 in hi:
            raise ValueError("ensembles take built-in boundary conditions only")
        self._lo, self._hi = lo, hi
        self._rp = s._rp_struct
        n = len(self.solutions)
        self._members = (L.EnsMember * n)()
        self._bufs = []
        ctl_size = ctypes.sizeof(L.StepCtl)
        self._ctl = torch.zeros(n * ctl_size, dtype=torch.uint8, device="cuda")
        raw = bytearray()
        for k, sol in enumerate(self.solutions):
            st = sol.state
            q0 = q_to_device(np.asarray(st.q), self._code)
            q1 = torch.empty_like(q0)
            comp = (torch.zeros_like(q0), torch.zeros_like(q0)) if self._code == L.F32K \
                else (None, None)
            self._bufs.append((q0, q1, comp[0], comp[1]))
            dx, dy = st.grid.delta
            m = self._members[k]
            m.mx, m.my = q0.shape[2], q0.shape[1]
            m.q0, m.q1 = q0.data_ptr(), q1.data_ptr()
            m.comp0 = comp[0].data_ptr() if comp[0] is not None else None
            m.comp1 = comp[1].data_ptr() if comp[1] is not None else None
            m.dx, m.dy = dx, dy
            m.ctl = self._ctl.data_ptr() + k * ctl_size
            c = L.StepCtl()
            c.t = sol.t
            c.dt = self.dts[k] if self.dts is not None else s.dt_initial
            c.tstart, c.tend = sol.t, float("inf")
            c.cfl = 0.0
            c.cfl_desired, c.cfl_max, c.dt_max = s.cfl_desired, s.cfl_max, s.dt_max
            c.cflmax, c.dtmin, c.dtmax = (s.status["cflmax"], s.status["dtmin"],
                                          s.status["dtmax"])
            c.dt_variable = int(bool(s.dt_variable))
            c.cur = 0
            c.max_steps = 2 ** 31 - 1
            raw += bytes(c)
        self._ctl.copy_(torch.frombuffer(raw, dtype=torch.uint8))
        self._prepare(self.rows)

    def _prepare(self, rows):
        n = len(self.solutions)
        nbytes = int(L.lib().clawhip_ensemble_table_bytes(n))
        if self._table is None or self._table.numel() < nbytes:
            self._table = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        waves = (ctypes.c_longlong * 3)()
        s = self.solver
        rc = L.lib().clawhip_ensemble_prepare(
            ctypes.byref(self._rp), self._code, self._members, n, L.ints(self._lo),
            L.ints(self._hi), L.ints(s.reflect_index), 2, L.ints(s._method),
            L.ints(list(s._mthlim) + [0] * (4 - len(s._mthlim))), int(rows),
            ctypes.c_void_p(self._table.data_ptr()), waves)
        L.check(rc, "clawhip_ensemble_prepare")
        self._waves = waves
        self.group_size = int(waves[2])
        self.rows = rows
        self._graphs = {}

    # ---------------------------------------------------------------- stepping
    def _launch(self, nsteps):
        rc = L.lib().clawhip_ensemble_run(ctypes.byref(self._rp), self._run_code,
                                          ctypes.c_void_p(self._table.data_ptr()), self._waves,
                                          int(nsteps), K._stream())
        L.check(rc, "clawhip_ensemble_run")

    def run_steps(self, nsteps, use_graph=True, chunk=16):
        """nsteps attempted steps of every member (no host synchronisation
        inside; whole chunks replay a captured graph)."""
        done = 0
        while done < nsteps:
            n = min(chunk, nsteps - done)
            if use_graph and n == chunk:
                g = self._graphs.get(n)
                if g is None:
                    from .device import no_gc
                    g = torch.cuda.CUDAGraph()
                    side = torch.cuda.Stream()
                    side.wait_stream(torch.cuda.current_stream())
                    with no_gc(), torch.cuda.stream(side):
                        with torch.cuda.graph(g, stream=side):
                            self._launch(n)
                    torch.cuda.current_stream().wait_stream(side)
                    self._graphs[n] = g
                g.replay()
            else:
                self._launch(n)
            done += n

    def autotune(self, candidates=(4, 6, 8, 10, 12, 16, 24, 32, 48, 64), steps=4, repeat=2,
                 group_sizes=(1, 2, 4)):
        """Times `steps` ensemble steps per candidate launch shape -- rows per
        wavefront, then wavefronts per workgroup at the best rows (the strip
        width is the 64 lanes of a wavefront) -- on copies of the members'
        controllers, so the ensemble's state does not advance, and keeps the
        fastest.  Returns {rows: seconds per step} (4 wavefronts per workgroup)
        and records {wavefronts per workgroup: seconds} in self.group_timing."""
        saved_ctl = self._ctl.clone()
        saved = [[t.clone() if t is not None else None for t in b[:4]] for b in self._bufs]

        def measure():
            best = None
            for _ in range(repeat):
                self._launch(1)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                self._launch(steps)
                torch.cuda.synchronize()
                el = (time.perf_counter() - t0) / steps
                best = el if best is None else min(best, el)
            self._ctl.copy_(saved_ctl)
            for b, sb in zip(self._bufs, saved):
                for t, st in zip(b[:4], sb):
                    if t is not None:
                        t.copy_(st)
            return best

        timing = {}
        for rows in candidates:
            self._prepare(rows)
            timing[rows] = measure()
        best_rows = min(timing, key=timing.get)
        self._prepare(best_rows)
        self.group_timing = {}
        for wpb in group_sizes:
            self._waves[2] = int(wpb)
            self.group_timing[int(wpb)] = measure()
        self.group_size = min(self.group_timing, key=self.group_timing.get)
        self._waves[2] = self.group_size
        self._graphs = {}
        return timing

    # ---------------------------------------------------------------- results
    def controllers(self):
        raw = self._ctl.cpu().numpy().tobytes()
        size = ctypes.sizeof(L.StepCtl)
        out = []
        for k in range(len(self.solutions)):
            c = L.StepCtl()
            ctypes_memmove(c, raw[k * size:(k + 1) * size])
            out.append(c)
        return out

    def member_q(self, k, controllers=None):
        """Host copy of member k's current state, (meqn, mx, my) as Solution.q
        (controllers: a controllers() list already read, to skip the read)."""
        c = (controllers if controllers is not None else self.controllers())[k]
        b = self._bufs[k]
        return q_from_device(b[c.cur] if c.cur in (0, 1) else b[0])

    def sync_solutions(self):
        """Write every member's state and time back to its Solution."""
        for k, (sol, c) in enumerate(zip(self.solutions, self.controllers())):
            b = self._bufs[k]
            sol.state.q = q_from_device(b[c.cur])
            sol.t = c.t

    @property
    def cells(self):
        return int(sum(m.mx * m.my for m in self._members))
