"""The ensemble (C4) and 3D split (C5) configurations against the oracle.

C4: members drawn from the ensemble workload's family (seed 4, sides
log-uniform in [64, 512], independent in x and y, seeded quadrant states, 2D
Euler Roe + MC unsplit with transverse correction, extrapolation BCs), advanced
by shared launches of Ensemble.run_steps, each compared with OracleRun stepping
that member alone (Solver.evolve_to_time one attempt at a time,
pyclaw/solver.py:554-666 with classic/step2.f90): FP64 bit for bit, FP32 + Kahan
within 2e-6 relative max per attempt.  One member starts with a dt above
cfl_max, so its first attempt is rejected and retried with the CFL's dt.

C5: the 3D Euler blast of the 3D split workload at 40 x 36 x 32 (unequal sides,
so a swapped axis shows), ClawSolver3D with FP32 + Kahan against OracleRun's
three step3ds sweeps (classic/step3ds.f90, classic/solver.py:648-706): relative
max within 2e-6 per step, accumulated over the steps, CFL within 1e-6, and the
same attempt rejected.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle.driver import OracleRun

pytestmark = pytest.mark.gpu

C4_MEMBERS = 8
C4_ATTEMPTS = 6
FP32K_BAR = 2e-6       # relative max error per attempt (FP32 + Kahan vs FP64)


def _c4_family(members, seed=4):
    """The ensemble workload's first `members` problems: sides and states."""
    from bench import quadrant_state
    rng = np.random.default_rng(seed)
    sides = np.rint(np.exp(rng.uniform(np.log(64), np.log(512), (512, 2)))).astype(int)
    return [((int(mx), int(my)), quadrant_state(int(mx), int(my), seed=1000 + k))
            for k, (mx, my) in enumerate(sides[:members])]


def _c4_solver(precision):
    import pyclaw_amd as pyclaw
    from pyclaw_amd import riemann
    s = pyclaw.ClawSolver2D(riemann.euler_4wave_2D)
    s.dimensional_split = False
    s.transverse_waves = s.trans_cor
    s.limiters = pyclaw.limiters.tvd.MC
    s.cfl_desired, s.cfl_max = 0.8, 1.0
    s.all_bcs = pyclaw.BC.extrap
    s.precision = precision
    s.dt_initial = 0.8 * (1.0 / 512) / 3.0
    return s


@pytest.mark.parametrize("precision", ["fp64", "fp64fast", "fp32k"])
def test_c4_ensemble_members_match_oracle(precision):
    import pyclaw_amd as pyclaw
    from pyclaw_amd.ensemble import Ensemble
    family = _c4_family(C4_MEMBERS)
    print("C4 ensemble vs oracle", precision, [s for s, _ in family], flush=True)
    sols = []
    for (mx, my), q in family:
        sol = pyclaw.Solution(4, pyclaw.Domain([0., 0.], [1., 1.], [mx, my]))
        sol.problem_data["gamma"] = 1.4
        sol.q = q
        sols.append(sol)
    solver = _c4_solver(precision)
    dts = [solver.dt_initial] * C4_MEMBERS
    dts[0] = 2.0 / family[0][0][0]          # Courant number ~3: rejected
    ens = Ensemble(solver, sols, dts=dts)
    runs = [OracleRun(O.make_rp(O.EULER, efix=1, gamma=1.4), q, [1.0 / mx, 1.0 / my],
                      [1, 1], [1, 1], dimensional_split=False, transverse=2, limiters=4,
                      cfl_desired=0.8, cfl_max=1.0, dt_initial=dts[k])
            for k, ((mx, my), q) in enumerate(family)]
    rejected = 0
    for attempt in range(1, C4_ATTEMPTS + 1):
        ens.run_steps(1, use_graph=False)
        ctls = ens.controllers()
        for k, run in enumerate(runs):
            n0 = run.numsteps
            run.evolve_to_time(None)
            rejected += run.numsteps == n0
            c = ctls[k]
            assert c.attempts == attempt and c.accepted == run.numsteps, (k, attempt)
            q = ens.member_q(k, ctls)
            if precision == "fp64":
                assert c.t == run.t and c.cfl == run.cfl, (k, attempt)
                assert np.array_equal(q, run.q), (k, attempt, np.abs(q - run.q).max())
            elif precision == "fp64fast":   # FMA-contracted FP64: 1e-12 per attempt
                assert abs(c.t - run.t) <= 1e-12 * run.t, (k, attempt)
                err = np.abs(q - run.q).max() / np.abs(run.q).max()
                assert err <= 1e-12 * attempt, (k, attempt, err)
            else:
                assert abs(c.t - run.t) <= 1e-6 * run.t, (k, attempt)
                err = np.abs(q - run.q).max() / np.abs(run.q).max()
                assert err <= FP32K_BAR * attempt, (k, attempt, err)
    assert rejected >= 1


def _blast(shape, seed=5):
    """The 3D split workload's seeded spherical blast on [-1, 1]^3."""
    rng = np.random.default_rng(seed)
    r0, pin = rng.uniform(0.1, 0.2), rng.uniform(5.0, 10.0)
    axes = [-1 + (np.arange(n) + 0.5) * (2.0 / n) for n in shape]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    p = np.where(X * X + Y * Y + Z * Z < (3 * r0) ** 2, pin, 0.1)
    q = np.zeros((5,) + tuple(shape), order="F")
    q[0] = 1.0 + 1e-3 * rng.uniform(-1, 1, X.shape)
    q[4] = p / 0.4
    return q, pin


def test_c5_split3_fp32k_matches_oracle():
    import pyclaw_amd as pyclaw
    from pyclaw_amd import riemann
    shape = (40, 36, 32)
    q, pin = _blast(shape)
    print("C5 fp32k split 3D vs oracle", shape, flush=True)
    solver = pyclaw.ClawSolver3D(riemann.euler_3D)
    solver.limiters = pyclaw.limiters.tvd.MC
    solver.all_bcs = pyclaw.BC.extrap
    solver.cfl_desired, solver.cfl_max = 0.8, 1.0
    solver.precision = "fp32k"
    sol = pyclaw.Solution(5, pyclaw.Domain([-1.] * 3, [1.] * 3, list(shape)))
    sol.problem_data["gamma"] = 1.4
    sol.q = q
    dt0 = 0.8 * (2.0 / max(shape)) / np.sqrt(1.4 * pin)
    solver.dt = solver.dt_initial = dt0
    solver.setup(sol)
    run = OracleRun(O.make_rp(O.EULER3, gamma=1.4), q, [2.0 / n for n in shape],
                    [1, 1, 1], [1, 1, 1], limiters=4, cfl_desired=0.8, cfl_max=1.0,
                    dt_initial=dt0)
    # the second attempt exceeds cfl_max as the blast accelerates: rejected, retried
    for attempt in range(1, 8):
        solver.evolve_to_time(sol)
        run.evolve_to_time(None)
        assert solver.status["numsteps"] == run.numsteps, attempt
        assert abs(sol.t - run.t) <= 1e-6 * run.t
        cfl = solver.cfl.get_cached_max()
        assert abs(cfl - run.cfl) <= 1e-6 * run.cfl, (attempt, cfl, run.cfl)
        err = np.abs(sol.state.q - run.q).max() / np.abs(run.q).max()
        assert err <= FP32K_BAR * run.numsteps, (attempt, err)
    assert run.numsteps == 6
