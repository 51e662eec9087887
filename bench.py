"""Benchmark: cell-updates/s of the PyClaw time-step hot path on one MI355X.

Headline workload (the metric's configuration): 2D shallow water (g = 1) on
the periodic unit square, 2048 x 2048 cells, SharpClaw characteristic WENO5 +
SSP33, CFL 0.45 (cfl_max 0.5).  Initial data: h = 1 + 8 Gaussian humps
(amplitude U[0.05, 0.3], width U[0.02, 0.1]) + 2 circular dam breaks (radius
U[0.05, 0.15], jump U[0.1, 0.5]) drawn from numpy.random.default_rng(3),
hu = hv = 0.  `value` is the FP32 + Kahan path with the learned troubled-
interface indicator (shipped weights; full WENO5 on flagged interfaces, the
linear weights elsewhere, full WENO5 judging every step's positivity); the
plain FP64 full-WENO5 path (bitwise equal to the reference step) runs the same
steps beside it, and their ratio and the error against it are reported.

A step is one full SSP33 time step (dt from the previous step's CFL,
accept/reject) on the state resident in HBM.  W untimed warm-up steps, then
exactly K timed steps.  The `full_run` block runs both paths to t = 1.0 and
reports the whole-run rate ratio, the error at t = 1.0 and the indicator's
trajectory.  Legs beside it: the classic 2D Euler step (1024^2), acoustics
(256^2), the 512-member ensemble and 3D Euler (384^3).

Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MX = MY = 1024              # classic Euler leg
SHARP_N = 2048              # the headline grid
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 vector
FP32_PEAK_TFLOPS = 157.3    # MI355X FP32 vector
BYTES_PER_CELL_F64 = 64     # classic Euler: read q + write q, 4 fields x 8 B
BYTES_PER_CELL_F32K = 64    # q + compensation, read + write, 4 fields x 4 B x 2
# SSP33 shallow water, algorithmic HBM bytes per cell-step (3 fields): FP64 reads
# q^n, q^1, q^n, q^2, q^n and writes q^1, q^2, q^n+1 (8 passes x 24 B); FP32 + Kahan
# the same 8 passes x 12 B plus the compensation read and written (2 x 12 B).
# One SSP33 step is 3 stage launches, so a launch carries a third of it.
SHARP_BYTES_F64 = 192
SHARP_BYTES_F32K = 120
SHARP_STAGES = 3


def quadrant_state(mx=MX, my=MY, seed=2, gamma=1.4):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.5, 1.5, 4)
    p = rng.uniform(0.1, 1.0, 4)
    u = rng.uniform(-0.5, 0.5, 4)
    v = rng.uniform(-0.5, 0.5, 4)
    xc = (np.arange(mx) + 0.5) / mx
    yc = (np.arange(my) + 0.5) / my
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    quad = (X >= 0.5).astype(int) + 2 * (Y >= 0.5).astype(int)
    q = np.empty((4, mx, my), order="F")
    q[0] = rho[quad]
    q[1] = rho[quad] * u[quad]
    q[2] = rho[quad] * v[quad]
    q[3] = p[quad] / (gamma - 1.0) + 0.5 * rho[quad] * (u[quad] ** 2 + v[quad] ** 2)
    return q


def make_problem(precision, mx=MX, my=MY):
    """ClawSolver2D + Solution of the headline workload."""
    import pyclaw_amd as pyclaw
    from pyclaw_amd import riemann
    solver = pyclaw.ClawSolver2D(riemann.euler_4wave_2D)
    solver.dimensional_split = False
    solver.transverse_waves = solver.trans_cor
    solver.limiters = pyclaw.limiters.tvd.MC
    solver.cfl_desired = 0.8
    solver.cfl_max = 1.0
    solver.all_bcs = pyclaw.BC.extrap
    solver.precision = precision
    domain = pyclaw.Domain([0., 0.], [1., 1.], [mx, my])
    solution = pyclaw.Solution(4, domain)
    solution.problem_data["gamma"] = 1.4
    solution.q = quadrant_state(mx, my)
    solver.dt = 0.8 * (1.0 / mx) / 3.0
    solver.dt_initial = solver.dt
    return solver, solution


class ChunkTimer(object):
    """HIP events around every replay of the captured 16-step graph, on the
    stream the graph runs on; kernel_ms = mean event time per step kernel."""

    def __init__(self, solver):
        import torch
        self.torch = torch
        self.solver = solver
        self.pairs = []
        self.orig = solver._replay_graph

        def timed(bufs, args, n):
            s = torch.cuda.current_stream()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(s)
            self.orig(bufs, args, n)
            e1.record(s)
            self.pairs.append((e0, e1, n))
        solver._replay_graph = timed

    def close(self):
        self.solver._replay_graph = self.orig

    def kernel_ms(self):
        self.torch.cuda.synchronize()
        tot = sum(a.elapsed_time(b) for a, b, _ in self.pairs)
        n = sum(k for _, _, k in self.pairs)
        return tot / max(n, 1), n


def run_gpu(precision, steps, warmup, use_graph=True):
    """At least one full graph chunk of untimed warm-up steps (so the graph is
    captured and instantiated before the timed region, whatever `warmup` is),
    then `steps` timed device-loop steps (graph replays of 16 steps)."""
    import torch
    solver, solution = make_problem(precision)
    solver.use_graph = use_graph
    solver.run_steps(solution, max(int(warmup), solver.graph_chunk))
    torch.cuda.synchronize()
    timer = ChunkTimer(solver)
    try:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        host = solver.run_steps(solution, steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    finally:
        timer.close()
    kms, nk = timer.kernel_ms()
    if nk == 0:          # eager launches: the wall time per step stands in
        kms = 1e3 * wall / steps
    return wall, kms, nk, host, solution


def measured_counts(name="k_unsplit_ctl_euler2d_1024"):
    """Per-launch counter figures of a kernel on its workload, from the
    rocprofv3 runs summarised in profiles/<name>.traffic.json: HBM bytes
    (FETCH_SIZE and WRITE_SIZE, each scaled by the same counter's reading for a
    256 MiB copy) and executed vector operations (SQ_INSTS_VALU_* x 64 lanes,
    FMA counted twice).  {} if absent."""
    path = os.path.join(ROOT, "profiles", name + ".traffic.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def classic_cpu_baseline(sample_steps=3):
    """The reference's own Fortran step2 (classic/step2.f90 + flux2.f90, built from
    the reference sources by oracle/ref.mk), single thread, on the same 1024^2
    state: `sample_steps` unsplit steps with the reference BC fill before each."""
    from oracle import oracle as O
    kind = "reference"
    if not O.ref_available():
        kind = "port"
    q = quadrant_state()
    ng = 2
    qbc = O.with_ghosts(q, ng)
    rp = O.make_rp(O.EULER, efix=1, gamma=1.4)
    meth = O.method_array(transverse=2)
    dx = 1.0 / MX
    dt = 0.8 * dx / 3.0
    t0 = time.perf_counter()
    for _ in range(sample_steps):
        O.bc2(qbc, ng, [1, 1], [1, 1])
        qnew = qbc.copy(order="F")
        if kind == "reference":
            O.ref_step2(rp, qbc, qnew, ng, dx, dx, dt, meth, [4] * 4)
        else:
            O.step2(rp, qbc, qnew, ng, dx, dx, dt, meth, [4] * 4)
        qbc = qnew
    wall = time.perf_counter() - t0
    return {"value": MX * MY * sample_steps / wall, "unit": "cell-updates/s", "cores": 1,
            "kind": kind,
            "sample": f"{sample_steps} unsplit Euler steps on the same 1024x1024 state "
                      f"(reference BC fill + classic/step2.f90 compiled with amdflang -O2), "
                      f"{wall:.1f} s"}


def acoustics_leg(n=256, seed=1, t_final=0.6):
    """2D acoustics (rho 1, bulk 4, c 2) on [0,1]^2, 256^2, solid walls, classic
    unsplit with transverse order 2, MC, CFL 0.9 (max 1.0), dt_initial 0.9 dx / c,
    to t = 0.6: p = exp(-100 r^2) + 1e-3 U[-1,1) (seed 1), u = v = 0.  FP64 and
    FP32 + Kahan through the device-controlled evolve_to_time.  The run to
    t = 0.6 (untimed: it includes the graph capture) gives the errors; the rate
    is timed on the continuation from t = 0.6 to 1.2."""
    import torch
    import pyclaw_amd as pyclaw
    from pyclaw_amd import riemann
    rng = np.random.default_rng(seed)
    xc = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(xc, xc, indexing="ij")
    q0 = np.zeros((3, n, n), order="F")
    q0[0] = np.exp(-100.0 * ((X - 0.5) ** 2 + (Y - 0.5) ** 2)) + 1e-3 * rng.uniform(-1, 1, X.shape)
    out, qs = {}, {}
    for prec in ("fp64", "fp32k"):
        solver = pyclaw.ClawSolver2D(riemann.acoustics_2D)
        solver.dimensional_split = False
        solver.transverse_waves = 2
        solver.limiters = pyclaw.limiters.tvd.MC
        solver.cfl_desired, solver.cfl_max = 0.9, 1.0
        solver.all_bcs = pyclaw.BC.wall
        solver.precision = prec
        sol = pyclaw.Solution(3, pyclaw.Domain([0., 0.], [1., 1.], [n, n]))
        sol.problem_data.update({"rho": 1.0, "bulk": 4.0, "cc": 2.0, "zz": 2.0})
        sol.q = q0
        solver.dt_initial = 0.9 * (1.0 / n) / 2.0
        solver.dt = solver.dt_initial
        solver.evolve_to_time(sol, t_final)
        steps_to_tfinal = solver.status["numsteps"]
        qs[prec] = sol.state.q.copy()
        t_at = sol.t
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solver.evolve_to_time(sol, 2 * t_final)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        steps = solver.status["numsteps"] - steps_to_tfinal
        out[prec] = {"value": n * n * steps / wall, "steps_to_t_final": steps_to_tfinal,
                     "t_final": t_at, "timed_steps": steps, "ms_per_step": 1e3 * wall / steps,
                     "hbm_frac": 48.0 * n * n / (wall / steps) / 1e9 / HBM_PEAK_GBS}
    qa, qb = qs["fp64"], qs["fp32k"]
    out["fp32k"]["rel_l1_vs_fp64"] = float(np.abs(qb - qa).sum() / np.abs(qa).sum())
    out["fp32k"]["rel_max_vs_fp64"] = float(np.abs(qb - qa).max() / np.abs(qa).max())
    out.update({"workload": "acoustics2d_gaussian_walls_unsplit_mc_256_t0.6",
                "unit": "cell-updates/s", "bytes_per_cell": 48})
    return out


def classic_roofline_leg(n=4096, steps=64, seed=1):
    """The plain FP64 classic step of the acoustics configuration (unsplit,
    transverse order 2, MC, solid walls, CFL 0.9) at 4096^2, where a step moves
    805 MB: the size at which the HBM roofline, not launch latency, bounds it.
    Device-controlled steps in captured graphs; HIP events around every replay
    (the step kernel and its one-wavefront control kernel)."""
    import torch
    import pyclaw_amd as pyclaw
    from pyclaw_amd import riemann
    rng = np.random.default_rng(seed)
    xc = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(xc, xc, indexing="ij")
    q0 = np.zeros((3, n, n), order="F")
    q0[0] = np.exp(-100.0 * ((X - 0.5) ** 2 + (Y - 0.5) ** 2)) + 1e-3 * rng.uniform(-1, 1, X.shape)
    del X, Y
    solver = pyclaw.ClawSolver2D(riemann.acoustics_2D)
    solver.dimensional_split = False
    solver.transverse_waves = 2
    solver.limiters = pyclaw.limiters.tvd.MC
    solver.cfl_desired, solver.cfl_max = 0.9, 1.0
    solver.all_bcs = pyclaw.BC.wall
    sol = pyclaw.Solution(3, pyclaw.Domain([0., 0.], [1., 1.], [n, n]))
    sol.problem_data.update({"rho": 1.0, "bulk": 4.0, "cc": 2.0, "zz": 2.0})
    sol.q = q0
    solver.dt_initial = 0.9 * (1.0 / n) / 2.0
    solver.dt = solver.dt_initial
    solver.run_steps(sol, solver.graph_chunk)
    torch.cuda.synchronize()
    timer = ChunkTimer(solver)
    try:
        t0 = time.perf_counter()
        solver.run_steps(sol, steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    finally:
        timer.close()
    kms, nk = timer.kernel_ms()
    achieved = 48.0 * n * n / (kms * 1e-3) / 1e9
    return {"workload": "acoustics2d_gaussian_walls_unsplit_mc_4096", "grid": [n, n],
            "value": n * n * steps / wall, "unit": "cell-updates/s", "steps": steps,
            "ms_per_step": 1e3 * wall / steps, "kernel_ms": kms, "kernels_timed": nk,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "kernel": "k_unsplit_ctl<Acoustics<double>> + k_step_ctl",
                         "bytes_per_cell": 48}}


def ensemble_leg(steps=32, members=512, seed=4):
    """512 independent 2D Euler Roe/MC unsplit problems (sides log-uniform in
    [64, 512], independent in x and y; seeded quadrant states as the headline
    workload), FP32 + Kahan, each with its own CFL-controlled dt, advanced by
    shared graph-captured launches; slices per wavefront autotuned."""
    import torch
    import pyclaw_amd as pyclaw
    from pyclaw_amd.ensemble import Ensemble
    rng = np.random.default_rng(seed)
    sides = np.rint(np.exp(rng.uniform(np.log(64), np.log(512), (members, 2)))).astype(int)
    sols = []
    for k, (mx, my) in enumerate(sides):
        sol = pyclaw.Solution(4, pyclaw.Domain([0., 0.], [1., 1.], [int(mx), int(my)]))
        sol.problem_data["gamma"] = 1.4
        sol.q = quadrant_state(int(mx), int(my), seed=1000 + k)
        sols.append(sol)
    solver, _ = make_problem("fp32k", 64, 64)
    solver.dt_initial = 0.8 * (1.0 / 512) / 3.0
    ens = Ensemble(solver, sols)
    timing = ens.autotune()
    ens.run_steps(16)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.run_steps(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cells = ens.cells
    acc = sum(c.accepted for c in ens.controllers())
    # the same members and steps in FP64 (untimed): error of the FP32 + Kahan run
    sols64 = []
    for k, (mx, my) in enumerate(sides):
        sol = pyclaw.Solution(4, pyclaw.Domain([0., 0.], [1., 1.], [int(mx), int(my)]))
        sol.problem_data["gamma"] = 1.4
        sol.q = quadrant_state(int(mx), int(my), seed=1000 + k)
        sols64.append(sol)
    solver64, _ = make_problem("fp64", 64, 64)
    solver64.dt_initial = solver.dt_initial
    ens64 = Ensemble(solver64, sols64)
    ens64.run_steps(16 + steps)
    diff = norm = mx_d = mx_n = 0.0
    ctl64, ctl32 = ens64.controllers(), ens.controllers()
    for k in range(members):
        qa, qb = ens64.member_q(k, ctl64), ens.member_q(k, ctl32)
        diff += float(np.abs(qb - qa).sum())
        norm += float(np.abs(qa).sum())
        mx_d, mx_n = max(mx_d, float(np.abs(qb - qa).max())), max(mx_n, float(np.abs(qa).max()))
    return {"workload": "ensemble_512_euler2d_roe_mc_unsplit_fp32k_loguniform_64_512",
            "rel_l1_vs_fp64": diff / norm, "rel_max_vs_fp64": mx_d / mx_n,
            "value": cells * steps / wall, "unit": "cell-updates/s", "members": members,
            "cells": cells, "steps": steps, "ms_per_step": 1e3 * wall / steps,
            "rows_per_wavefront": ens.rows,
            "wavefronts_per_workgroup": ens.group_size,
            "autotune_ms": {str(k): 1e3 * v for k, v in timing.items()},
            "autotune_workgroup_ms": {str(k): 1e3 * v for k, v in ens.group_timing.items()},
            "accepted_fraction": acc / float(members * (steps + 16)),
            "hbm_frac": BYTES_PER_CELL_F32K * cells / (wall / steps) / 1e9 / HBM_PEAK_GBS}


def euler3d_leg(steps=10, n=384, seed=5):
    """3D Euler, Roe, dimension-split, MC, 384^3, extrapolation BCs: seeded
    spherical blast (radius U[0.1,0.2], inner p U[5,10], outer p 0.1, rho 1 with
    1e-3 seeded noise) on [-1,1]^3; FP64, fp64fast and FP32 + Kahan."""
    import torch
    import pyclaw_amd as pyclaw
    from pyclaw_amd import riemann
    rng = np.random.default_rng(seed)
    r0, pin = rng.uniform(0.1, 0.2), rng.uniform(5.0, 10.0)
    x = -1 + (np.arange(n) + 0.5) * (2.0 / n)
    out = {}
    q = None
    for prec in ("fp64", "fp64fast", "fp32k"):
        solver = pyclaw.ClawSolver3D(riemann.euler_3D)
        solver.limiters = pyclaw.limiters.tvd.MC
        solver.all_bcs = pyclaw.BC.extrap
        solver.cfl_desired, solver.cfl_max = 0.8, 1.0
        solver.precision = prec
        sol = pyclaw.Solution(5, pyclaw.Domain([-1.] * 3, [1.] * 3, [n] * 3))
        sol.problem_data["gamma"] = 1.4
        if q is None:
            X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
            p = np.where(X * X + Y * Y + Z * Z < r0 * r0, pin, 0.1)
            rho = 1.0 + 1e-3 * rng.uniform(-1, 1, X.shape)
            q = np.zeros((5, n, n, n), order="F")
            q[0] = rho
            q[4] = p / 0.4
            del X, Y, Z, p, rho
        sol.q = q
        solver.dt = 0.8 * (2.0 / n) / np.sqrt(1.4 * pin)
        solver.setup(sol)
        solver.evolve_to_time(sol)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            solver.evolve_to_time(sol)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        out[prec] = {"value": n ** 3 * steps / wall, "ms_per_step": 1e3 * wall / steps,
                     "hbm_frac": 240.0 * n ** 3 / (wall / steps) / 1e9 / HBM_PEAK_GBS}
        out[prec + "_q"] = sol.state.q
    qa = out.pop("fp64_q")
    for prec in ("fp64fast", "fp32k"):
        qb = out.pop(prec + "_q")
        out[prec]["rel_l1_vs_fp64"] = float(np.abs(qb - qa).sum() / np.abs(qa).sum())
        out[prec]["rel_max_vs_fp64"] = float(np.abs(qb - qa).max() / np.abs(qa).max())
    out.update({"workload": "euler3d_blast_roe_mc_split_384cubed", "unit": "cell-updates/s",
                "steps": steps, "bytes_per_cell_step": 240})
    return out


def sharp_problem(prec, indicator, n=SHARP_N, seed=3):
    from pyclaw_amd import indicator as I
    solver = I.swe_solver(prec, indicator)
    sol = I.swe_solution(seed, n, n)
    solver.dt = 0.45 / n / 1.6
    solver.setup(sol)
    return solver, sol


class ReplayTimer(object):
    """HIP events around every launch of device-loop attempts (graph replays,
    and the eager launches of a remainder), on the stream they run on."""

    def __init__(self):
        import torch
        from pyclaw_amd import kernels
        self.torch, self.kernels = torch, kernels
        self.pairs = []
        self.orig_run = kernels.sharp2_ssp33_run

        def timed_run(*a, **kw):
            if torch.cuda.is_current_stream_capturing():
                self.orig_run(*a, **kw)
            else:
                self._mark(lambda: self.orig_run(*a, **kw))
        kernels.sharp2_ssp33_run = timed_run
        self.graph_cls = torch.cuda.CUDAGraph
        self.orig_replay = torch.cuda.CUDAGraph.replay
        timer = self

        def timed_replay(g):
            timer._mark(lambda: timer.orig_replay(g))
        torch.cuda.CUDAGraph.replay = timed_replay

    def _mark(self, fn):
        st = self.torch.cuda.current_stream()
        e0 = self.torch.cuda.Event(enable_timing=True)
        e1 = self.torch.cuda.Event(enable_timing=True)
        e0.record(st)
        fn()
        e1.record(st)
        self.pairs.append((e0, e1))

    def close(self):
        self.kernels.sharp2_ssp33_run = self.orig_run
        self.graph_cls.replay = self.orig_replay

    def total_ms(self):
        self.torch.cuda.synchronize()
        return sum(a.elapsed_time(b) for a, b in self.pairs)


USE_GRAPH = True   # --no-graph: launch every attempt eagerly (per-dispatch counter runs)


def sharp_steps(prec, use_indicator, steps, warmup, n=SHARP_N):
    """`warmup` untimed attempts, the capture of the graphs the timed attempts
    replay, then exactly `steps` timed SSP33 attempts of the 2048^2
    shallow-water workload in the device-controlled loop (graph-captured chunks
    of 16 attempts, indicator audits on the host path).  HIP events around
    every replay give the GPU time of the timed attempts."""
    import torch
    from pyclaw_amd import indicator as I
    ind = I.TroubledInterfaceIndicator() if use_indicator else None
    solver, sol = sharp_problem(prec, ind, n)
    solver.use_graph = USE_GRAPH
    if warmup > 0:
        solver.run_steps(sol, warmup)
    solver.capture_graphs(sol, steps)
    torch.cuda.synchronize()
    acc0 = solver.status["numsteps"]
    timer = ReplayTimer()
    try:
        t0 = time.perf_counter()
        solver.run_steps(sol, steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    finally:
        timer.close()
    gpu_ms = timer.total_ms()
    # the dominant kernel, the stage march, timed on its own: 8 more attempts
    # launched without graphs, HIP events around each stage's march
    marches = solver.time_stages(sol, 8)
    stage_ms = float(np.mean(marches))
    per_cell = SHARP_BYTES_F32K if prec == "fp32k" else SHARP_BYTES_F64
    achieved = per_cell / SHARP_STAGES * n * n / (stage_ms * 1e-3) / 1e9
    out = {"value": n * n * steps / wall, "ms_per_step": 1e3 * wall / steps,
           "gpu_ms_per_step": gpu_ms / steps, "stage_ms": stage_ms,
           "stage_launches_timed": len(marches), "t_end": sol.t,
           "accepted": solver.status["numsteps"] - acc0,
           "hbm": {"achieved": achieved, "frac": achieved / HBM_PEAK_GBS,
                   "bytes_per_cell_step": per_cell}}
    if ind is not None:
        out["troubled_fraction"] = ind.troubled_fraction()
        out["flag_refresh_every"] = ind.refresh_interval(solver.cfl_max)
        out["judge"] = dict(ind.judge_stats)
    return out, sol


def rel_err(q, ref):
    return (float(np.abs(q - ref).sum() / np.abs(ref).sum()),
            float(np.abs(q - ref).max() / np.abs(ref).max()))


def full_run(n=SHARP_N, t_final=1.0, nout=10):
    """Both paths from t = 0 to t_final with output times every t_final/nout
    (the same clipped dt sequence for both): wall time, steps, the error of the
    FP32 + indicator state against the FP64 full-WENO5 state at every output
    time, and the indicator's troubled fraction and judge statistics."""
    import torch
    from pyclaw_amd import indicator as I
    res, frames = {}, {}
    for name, prec, use in (("fp64_weno", "fp64", False), ("fp32k_mlp", "fp32k", True)):
        ind = I.TroubledInterfaceIndicator() if use else None
        solver, sol = sharp_problem(prec, ind, n)
        traj, qs, wall = [], [], 0.0
        for t in np.linspace(0.0, t_final, nout + 1)[1:]:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            solver.evolve_to_time(sol, float(t))
            torch.cuda.synchronize()
            wall += time.perf_counter() - t0
            qs.append(sol.state.q.copy())
            if ind is not None:
                traj.append(ind.troubled_fraction())
        steps = solver.status["numsteps"]
        res[name] = {"wall_s": wall, "steps": steps, "t_end": sol.t,
                     "value": n * n * steps / wall}
        if ind is not None:
            res[name]["troubled_fraction_at_outputs"] = traj
            res[name]["judge"] = dict(ind.judge_stats)
        frames[name] = qs
        print(f"[bench] full run {name}: {steps} steps in {wall:.1f} s", file=sys.stderr,
              flush=True)
    errs = [rel_err(a, b) for a, b in zip(frames["fp32k_mlp"], frames["fp64_weno"])]
    res["rel_l1_vs_fp64_at_outputs"] = [e[0] for e in errs]
    res["rel_max_vs_fp64_at_outputs"] = [e[1] for e in errs]
    res["rel_l1_vs_fp64_at_t_final"], res["rel_max_vs_fp64_at_t_final"] = errs[-1]
    res["speedup_fp32k_mlp_vs_fp64_weno"] = res["fp32k_mlp"]["value"] / res["fp64_weno"]["value"]
    res.update({"t_final": t_final, "output_times": nout, "grid": [n, n]})
    return res


def sharp_cpu_baseline(sample_steps=2, n=SHARP_N, seed=3):
    """The reference's own SharpClaw operator (sharpclaw/flux2.f90 + flux1.f90 +
    reconstruct.f90 weno5_char, built from the reference sources by oracle/ref.mk)
    in the reference's SSP33 step (sharpclaw/solver.py:289-323, periodic ghost
    fill before each of the 3 operator evaluations, numpy stage combinations,
    dt from the CFL), single thread, on the same 2048^2 initial state."""
    from oracle import oracle as O
    from pyclaw_amd import indicator as I
    kind = "reference" if O.ref_available() else "port"
    flux = O.ref_sharp_flux2 if kind == "reference" else O.sharp_flux2
    rp = O.make_rp(O.SHALLOW, efix=1, grav=1.0)
    q = I.swe_problem(seed, n, n)
    dx = 1.0 / n
    dt = 0.45 / n / 1.6
    ng = 3

    def dq(y):
        ybc = O.with_ghosts(y, ng)
        O.bc2(ybc, ng, [2, 2], [2, 2])
        d, c = flux(rp, ybc, ng, dx, dx, dt, 2)
        return O.interior(d, ng), c

    t0 = time.perf_counter()
    for _ in range(sample_steps):
        d1, c1 = dq(q)
        dt_new = dt * 0.45 / c1
        d1 = d1 * (dt_new / dt)    # dq_dt*dt at the new dt
        dt = dt_new
        r = q + d1
        d2, _ = dq(r)
        r = 0.75 * q + 0.25 * (r + d2)
        d3, _ = dq(r)
        q = 1. / 3. * q + 2. / 3. * (r + d3)
    wall = time.perf_counter() - t0
    return {"value": n * n * sample_steps / wall, "unit": "cell-updates/s", "cores": 1,
            "kind": kind,
            "sample": f"{sample_steps} SSP33 steps of the {n}x{n} shallow-water workload "
                      f"(the reference's sharpclaw/flux2.f90 characteristic WENO5 compiled "
                      f"with amdflang -O2, periodic ghost fill and numpy stage updates), "
                      f"{wall:.1f} s"}


def _sharp_cpu_strip(rows, sample_steps, seed, n=SHARP_N):
    """One worker of the all-cores CPU baseline: the reference SharpClaw SSP33 step
    (as sharp_cpu_baseline) on an n x rows periodic strip of the shallow-water
    family; returns the compute wall time."""
    from oracle import oracle as O
    from pyclaw_amd import indicator as I
    flux = O.ref_sharp_flux2 if O.ref_available() else O.sharp_flux2
    rp = O.make_rp(O.SHALLOW, efix=1, grav=1.0)
    q = I.swe_problem(seed, n, rows)
    dx, dy = 1.0 / n, 1.0 / rows
    dt = 0.45 / n / 1.6
    ng = 3

    def dq(y):
        ybc = O.with_ghosts(y, ng)
        O.bc2(ybc, ng, [2, 2], [2, 2])
        d, c = flux(rp, ybc, ng, dx, dy, dt, 2)
        return O.interior(d, ng), c

    t0 = time.perf_counter()
    for _ in range(sample_steps):
        d1, c1 = dq(q)
        dt_new = dt * 0.45 / c1
        d1 = d1 * (dt_new / dt)
        dt = dt_new
        r = q + d1
        d2, _ = dq(r)
        r = 0.75 * q + 0.25 * (r + d2)
        d3, _ = dq(r)
        q = 1. / 3. * q + 2. / 3. * (r + d3)
    return time.perf_counter() - t0


def sharp_cpu_baseline_parallel(workers=16, sample_steps=4, n=SHARP_N, seed=3):
    """The same reference SharpClaw step on every core of the GPU box's CPU share:
    `workers` processes (child processes of this one, started before they are
    timed), each on an n x (n / workers) periodic strip, so that together they
    advance n^2 cells; they start together (a line on stdin) and the rate is the
    total over the slowest worker's wall time."""
    import subprocess
    rows = n // workers
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker",
                               str(rows), str(sample_steps), str(seed + k), str(n)],
                              stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True)
             for k in range(workers)]
    for p in procs:
        assert p.stdout.readline().strip() == "ready"
    for p in procs:
        p.stdin.write("go\n")
        p.stdin.flush()
    walls = []
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError("CPU baseline worker failed")
        walls.append(float(out.strip().splitlines()[-1]))
    wall = max(walls)
    from oracle import oracle as O
    return {"value": n * rows * workers * sample_steps / wall, "unit": "cell-updates/s",
            "cores": workers, "kind": "reference" if O.ref_available() else "port",
            "sample": f"{workers} processes, each {sample_steps} SSP33 steps of an {n}x{rows} "
                      f"periodic strip of the shallow-water family ({n}x{rows * workers} cells "
                      f"together; the reference's sharpclaw/flux2.f90 characteristic WENO5 "
                      f"compiled with amdflang -O2), slowest worker {wall:.1f} s"}


def _cpu_worker(argv):
    rows, steps, seed = int(argv[0]), int(argv[1]), int(argv[2])
    # imports and the reference build's loading happen before the start line
    from oracle import oracle as O
    from pyclaw_amd import indicator  # noqa: F401 (the problem generator)
    O.ref_available()
    print("ready", flush=True)
    sys.stdin.readline()
    print(f"{_sharp_cpu_strip(rows, steps, seed, n=int(argv[3])):.6f}", flush=True)


def classic_leg(steps=400, warmup=16):
    """The classic unsplit 2D Euler step on 1024^2 (FP64 and FP32 + Kahan),
    device-controlled, graph captured before the timed region."""
    cells = MX * MY
    wall64, k64, n64, host64, sol64 = run_gpu("fp64", steps, warmup, USE_GRAPH)
    wall32, k32, n32, host32, sol32 = run_gpu("fp32k", steps, warmup, USE_GRAPH)
    wallf, kf, nf, hostf, solf = run_gpu("fp64fast", steps, warmup, USE_GRAPH)
    err_l1, err_max = rel_err(sol32.state.q, sol64.state.q)
    errf_l1, errf_max = rel_err(solf.state.q, sol64.state.q)
    achieved = BYTES_PER_CELL_F64 * cells / (k64 * 1e-3) / 1e9
    counts = measured_counts()
    flops = counts.get("fp64_flops_per_launch")
    return {
        "workload": "euler2d_quadrants_roe_efix_unsplit_transcor_mc_1024x1024",
        "value": cells * steps / wall64, "unit": "cell-updates/s", "steps": steps,
        "ms_per_step": 1e3 * wall64 / steps, "t_end": sol64.t, "accepted": host64.accepted,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": counts.get("bytes_per_launch"),
                     "kernel": "k_unsplit_ctl<Euler<double>>", "kernel_ms": k64,
                     "kernels_timed": n64, "bytes_per_cell": BYTES_PER_CELL_F64},
        "fp64_valu": (None if flops is None else
                      {"achieved": flops / (k64 * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS,
                       "unit": "TFLOP/s",
                       "frac": flops / (k64 * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                       "flops_per_launch": flops}),
        "fp64_fast": {"value": cells * steps / wallf, "ms_per_step": 1e3 * wallf / steps,
                      "kernel_ms": kf,
                      "hbm_frac": BYTES_PER_CELL_F64 * cells / (kf * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "rel_l1_vs_fp64": errf_l1, "rel_max_vs_fp64": errf_max, "t_end": solf.t},
        "fp32_kahan": {"value": cells * steps / wall32, "ms_per_step": 1e3 * wall32 / steps,
                       "kernel_ms": k32,
                       "hbm_frac": BYTES_PER_CELL_F32K * cells / (k32 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                       "rel_l1_vs_fp64": err_l1, "rel_max_vs_fp64": err_max, "t_end": sol32.t},
        "cpu_baseline": classic_cpu_baseline(),
    }


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--cpu-worker":
        _cpu_worker(sys.argv[2:])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-full-run", action="store_true",
                    help="skip the runs to t = 1.0 of both SharpClaw paths")
    ap.add_argument("--extras", default="classic,acoustics,classic_roofline,ensemble,euler3d",
                    help="comma list of extra workloads to time beside the headline")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the device-loop attempts one by one (per-dispatch counters)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="end with a plain 256 MiB device copy (known bytes for counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the hot path runs on a single MI355X in this build "
              "(multi-GPU sharding is described in DESIGN.md, not run).", file=sys.stderr)
        sys.exit(2)
    global USE_GRAPH
    USE_GRAPH = not args.no_graph
    n = SHARP_N
    head, sol_mlp = sharp_steps("fp32k", True, args.steps, args.warmup)
    print("[bench] fp32k + indicator done", file=sys.stderr, flush=True)
    f64, sol64 = sharp_steps("fp64", False, args.steps, args.warmup)
    print("[bench] fp64 full WENO5 done", file=sys.stderr, flush=True)
    f32, sol32 = sharp_steps("fp32k", False, args.steps, args.warmup)
    print("[bench] fp32k full WENO5 done", file=sys.stderr, flush=True)
    f64f, sol64f = sharp_steps("fp64fast", False, args.steps, args.warmup)
    print("[bench] fp64fast full WENO5 done", file=sys.stderr, flush=True)
    f64f["rel_l1_vs_fp64"], f64f["rel_max_vs_fp64"] = rel_err(sol64f.state.q, sol64.state.q)
    l1, mx_err = rel_err(sol_mlp.state.q, sol64.state.q)
    f32["rel_l1_vs_fp64"], f32["rel_max_vs_fp64"] = rel_err(sol32.state.q, sol64.state.q)
    counts = measured_counts("sharp_swe2d_2048_fp32k_mlp")
    rec = {
        "metric": "cell-updates/s, 2D Euler/SWE 2048\u00b2: FP64 plain vs FP32+learned-WENO; "
                  "% HBM peak",
        "value": head["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "swe2d_periodic_humps_dambreaks_sharpclaw_weno5char_ssp33_2048",
                   "grid": [n, n], "cfl": 0.45, "cfl_max": 0.5, "bc": "periodic", "seed": 3,
                   "time_integrator": "SSP33", "char_decomp": 2,
                   "precision": "fp32 state and fluxes, Kahan-compensated update, fp64 CFL",
                   "indicator": "shipped weights, bf16 MFMA, flags every 2nd step",
                   "t_end": head["t_end"], "accepted": head["accepted"]},
        "roofline": {"bound": "hbm", "achieved": head["hbm"]["achieved"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": head["hbm"]["frac"],
                     "traffic": counts.get("bytes_per_launch"),
                     "kernel": "k_sharp_march<Shallow<float>, float, true, true>",
                     "kernel_ms": head["stage_ms"],
                     "kernels_timed": head["stage_launches_timed"],
                     "bytes_per_cell_per_launch": SHARP_BYTES_F32K / SHARP_STAGES},
        "indicator": {k: head[k] for k in ("troubled_fraction", "flag_refresh_every", "judge")},
        "fp64_weno": f64,
        "fp32k_weno": f32,
        "fp64fast_weno": f64f,
        "speedup_vs_fp64_weno": head["value"] / f64["value"],
        "rel_l1_vs_fp64": l1,
        "rel_max_vs_fp64": mx_err,
    }
    if not args.no_full_run:
        rec["full_run"] = full_run()
        print("[bench] full run done", file=sys.stderr, flush=True)
    legs = {"classic": classic_leg, "acoustics": acoustics_leg, "ensemble": ensemble_leg,
            "euler3d": euler3d_leg, "classic_roofline": classic_roofline_leg}
    for name in [x for x in args.extras.split(",") if x]:
        try:
            rec[name] = legs[name]()
        except Exception as e:   # an extra leg never hides the headline line
            rec[name] = {"error": repr(e)}
        print(f"[bench] {name} done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        rec["cpu_baseline"] = sharp_cpu_baseline()
        try:
            rec["cpu_baseline_all_cores"] = sharp_cpu_baseline_parallel()
        except Exception as e:   # never hides the headline line
            rec["cpu_baseline_all_cores"] = {"error": repr(e)}
    if args.calibrate_copy:
        import torch
        a = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
