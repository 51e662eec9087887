"""FP64 reductions over a state (stats.hip: conserved totals, positivity counts,
L1 distance) against numpy, conservation of the totals by the SharpClaw steps,
and full WENO5 judging the learned indicator's steps (SharpClawSolver._judge)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("shape", [(3, 37, 53), (4, 64, 64), (5, 9, 11, 13)])
def test_state_sums_match_numpy(dtype, shape):
    from pyclaw_amd import kernels as K
    rng = np.random.default_rng(sum(shape))
    a = rng.uniform(0.5, 2.0, shape)
    a[0].flat[rng.integers(0, a[0].size, 7)] = -1.0      # non-positive cells in field 0
    a[1].flat[rng.integers(0, a[1].size, 5)] = 0.0
    a[2].flat[3] = np.nan                                 # not finite: counted, sum NaN
    q = torch.tensor(a, dtype=dtype, device="cuda")
    out = K.state_sums(q, positive=(0, 1)).cpu().numpy()
    host = q.cpu().double().numpy()
    for f in range(shape[0]):
        ref = host[f].sum()
        if f == 2:
            assert np.isnan(out[f])
        else:
            assert abs(out[f] - ref) <= 1e-12 * abs(host[f]).sum()
    bad = (~(host[0] > 0)) | (~(host[1] > 0))
    assert out[-1] == bad.sum()
    # no mask: nothing counted; deterministic order: a second call is bitwise equal
    out2 = K.state_sums(q).cpu().numpy()
    assert out2[-1] == 0
    np.testing.assert_array_equal(out2[:-1], out[:-1])


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_l1_diff_matches_numpy(dtype):
    from pyclaw_amd import kernels as K
    rng = np.random.default_rng(5)
    a = torch.tensor(rng.normal(size=(3, 300, 257)), dtype=dtype, device="cuda")
    b = torch.tensor(rng.normal(size=(3, 300, 257)), dtype=dtype, device="cuda")
    d, n = K.l1_diff(a, b).cpu().numpy()
    ha, hb = a.cpu().double().numpy(), b.cpu().double().numpy()
    assert abs(d - np.abs(ha - hb).sum()) <= 1e-12 * d
    assert abs(n - np.abs(hb).sum()) <= 1e-12 * n
    assert K.l1_diff(a, a).cpu().numpy()[0] == 0.0


def _totals(sol, solver):
    from pyclaw_amd import kernels as K
    q, _ = sol.state.device_q(solver._dtype_code)
    return K.state_sums(q).cpu().numpy()[:-1]


@pytest.mark.parametrize("prec,ind", [("fp64", False), ("fp32k", False), ("fp32k", True)])
def test_sharpclaw_periodic_conserves_totals(prec, ind):
    """Periodic shallow water: the FP64 totals of h, hu, hv stay at their initial
    values to round-off (the flux-difference update telescopes)."""
    from pyclaw_amd import indicator as I
    n = 128
    solver = I.swe_solver(prec, I.TroubledInterfaceIndicator() if ind else None)
    sol = I.swe_solution(7, n, n)
    solver.dt = 0.45 / n / 1.6
    solver.setup(sol)
    solver.evolve_to_time(sol)
    t0 = _totals(sol, solver)
    for _ in range(40):
        solver.evolve_to_time(sol)
    t1 = _totals(sol, solver)
    scale = n * n       # h ~ 1 per cell; momentum totals start at 0
    tol = 1e-13 if prec == "fp64" else 2e-6
    assert np.all(np.abs(t1 - t0) <= tol * scale), (t0, t1)


def _one_step(prec, indicator=None, seed=3, n=192):
    from pyclaw_amd import indicator as I
    solver = I.swe_solver(prec, indicator)
    sol = I.swe_solution(seed, n, n)
    solver.dt = 0.45 / n / 1.6
    solver.setup(sol)
    solver.evolve_to_time(sol)
    return sol.state.q.copy(), sol.t, solver


def test_failed_audit_keeps_the_full_weno_step():
    """audit_every = 1 with a tolerance no step can meet: the step's result is the
    full-WENO5 step, bit for bit, and the flags switch to every-step refresh."""
    from pyclaw_amd import indicator as I
    ind = I.TroubledInterfaceIndicator()
    ind.audit_every, ind.audit_tol = 1, -1.0
    qa, ta, _ = _one_step("fp32k", ind)
    qb, tb, _ = _one_step("fp32k", None)
    assert ta == tb
    np.testing.assert_array_equal(qa, qb)
    assert ind.judge_stats["audits"] == 1 and ind.judge_stats["audit_failures"] == 1
    assert ind.every == 1


def test_passing_audits_keep_the_indicator_step():
    from pyclaw_amd import indicator as I
    ind = I.TroubledInterfaceIndicator()
    ind.audit_every = 2
    solver = I.swe_solver("fp32k", ind)
    n = 192
    sol = I.swe_solution(3, n, n)
    solver.dt = 0.45 / n / 1.6
    solver.setup(sol)
    for _ in range(10):
        solver.evolve_to_time(sol)
    st = ind.judge_stats
    print(st)
    assert st["steps"] == 10 and st["audits"] == 5
    assert st["audit_failures"] == 0 and st["positivity_fallbacks"] == 0
    assert 0 < st["max_audit_rel_l1"] <= ind.audit_tol


def test_positivity_failure_recomputes_with_full_weno():
    """A step whose result fails the positivity check (here: a field that is zero
    in places, declared positive) is recomputed with full WENO5."""
    from pyclaw_amd import indicator as I
    ind = I.TroubledInterfaceIndicator()
    ind.audit_every = 0
    from pyclaw_amd import SharpClawSolver2D
    orig = SharpClawSolver2D._positive_fields
    SharpClawSolver2D._positive_fields = lambda self: (1,)   # hu: zero / negative somewhere
    try:
        qa, ta, _ = _one_step("fp32k", ind)
    finally:
        SharpClawSolver2D._positive_fields = orig
    qb, tb, _ = _one_step("fp32k", None)
    assert ta == tb
    np.testing.assert_array_equal(qa, qb)
    assert ind.judge_stats["positivity_fallbacks"] == 1
