"""Ensembles (clawhip_ensemble_run): every member of an uneven-size ensemble
advanced by shared launches equals the same problem stepped alone by
ClawSolver2D.run_steps -- bitwise, with the same t after the same attempts."""
import numpy as np
import pytest

from oracle import oracle as O
from tests._cases import seeded_state

pytestmark = pytest.mark.gpu

SIZES = [(37, 53), (64, 64), (130, 71), (9, 200), (61, 61), (121, 120)]


def _solver(precision):
    import pyclaw_amd as pyclaw
    from pyclaw_amd import riemann
    s = pyclaw.ClawSolver2D(riemann.euler_4wave_2D)
    s.dimensional_split = False
    s.transverse_waves = s.trans_cor
    s.limiters = pyclaw.limiters.tvd.MC
    s.cfl_desired, s.cfl_max = 0.8, 1.0
    s.bc_lower[0], s.bc_upper[0] = pyclaw.BC.extrap, pyclaw.BC.wall
    s.bc_lower[1], s.bc_upper[1] = pyclaw.BC.wall, pyclaw.BC.extrap
    s.precision = precision
    s.dt_initial = 0.002
    return s


def _solution(mx, my, seed):
    import pyclaw_amd as pyclaw
    sol = pyclaw.Solution(4, pyclaw.Domain([0., 0.], [1., 1.2], [mx, my]))
    sol.problem_data["gamma"] = 1.4
    sol.q = seeded_state(np.random.default_rng(seed), O.EULER, (mx, my))
    return sol


@pytest.mark.parametrize("precision", ["fp64", "fp64fast", "fp32k"])
def test_ensemble_members_equal_single_runs(precision):
    from pyclaw_amd.ensemble import Ensemble
    sols = [_solution(mx, my, 100 + k) for k, (mx, my) in enumerate(SIZES)]
    ens = Ensemble(_solver(precision), sols)
    timing = ens.autotune(candidates=(4, 8, 16), steps=2, repeat=1)
    assert set(timing) == {4, 8, 16}
    assert set(ens.group_timing) == {1, 2, 4} and ens.group_size in (1, 2, 4)
    assert all(c.attempts == 0 for c in ens.controllers())   # tuning did not advance
    ens.run_steps(21)
    ctls = ens.controllers()
    for k, (mx, my) in enumerate(SIZES):
        solver = _solver(precision)
        sol = _solution(mx, my, 100 + k)
        solver.dt = solver.dt_initial
        host = solver.run_steps(sol, 21)
        assert ctls[k].attempts == 21 == host.attempts
        assert ctls[k].accepted == host.accepted and ctls[k].t == host.t, k
        assert np.array_equal(ens.member_q(k), sol.state.q), (k, (mx, my))
