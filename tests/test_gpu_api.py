"""GPU: dimension-split steps bit for bit, and whole runs through the PyClaw API.

The end-to-end cases are the reference's own example set-ups (examples/*/); each
runs through pyclaw_amd's Controller/Solver on the GPU and must
  * meet the reference test's tolerance against the reference fixture, and
  * equal the oracle's run of the same set-up bit for bit (same dt sequence:
    same number of accepted steps, same final time, same state).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from oracle.driver import OracleRun
from tests import _cases as C
from tests._cases import seeded_state

pytestmark = pytest.mark.gpu


def _pyclaw():
    import pyclaw_amd
    return pyclaw_amd


# ---------------------------------------------------------------- single steps
@pytest.mark.parametrize("kind", [O.ACOUSTICS, O.EULER, O.SHALLOW, O.ADVECTION])
def test_split2_step_bit_exact(kind):
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device, q_to_device
    from tests.test_gpu_classic2 import KIND_PD, _oracle_rp
    rng = np.random.default_rng(40 + kind)
    mw = O.MWAVES[kind]
    for (mx, my), (bcl, bcu) in zip([(67, 35), (130, 70), (4, 3)],
                                    [([1, 3], [2, 2]), ([2, 1], [2, 3]), ([3, 3], [1, 1])]):
        if kind == O.ADVECTION:
            bcl, bcu = [2, 1], [2, 1]
        q = seeded_state(rng, kind, (mx, my))
        meth = O.method_array(dimensional_split=True)
        dx, dy = 1.0 / mx, 0.7 / my
        dt = 0.25 * min(dx, dy)
        ng = 2
        qbc = O.with_ghosts(q, ng)
        O.bc2(qbc, ng, bcl, bcu)
        qold = qbc.copy(order="F")
        c1 = O.step2(_oracle_rp(kind), qold, qbc, ng, dx, dy, dt, meth, [4] * mw, ids=1)
        c2 = O.step2(_oracle_rp(kind), qbc, qbc, ng, dx, dy, dt, meth, [4] * mw, ids=2)
        ref = O.interior(qbc, ng)
        qi = q_to_device(q)
        qo = torch.empty_like(qi)
        cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
        K.step2_split(K.make_rp(kind, KIND_PD[kind]), qi, qo, cfl, bcl, bcu, [1, 2], meth,
                      [4] * mw, dx, dy, dt)
        torch.cuda.synchronize()
        out = q_from_device(qo)
        assert np.array_equal(out, ref), (mx, my, np.abs(out - ref).max())
        assert float(cfl.item()) == max(c1, c2)


def test_split3_step_bit_exact():
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device, q_to_device
    rng = np.random.default_rng(3)
    for (mx, my, mz), (bcl, bcu) in zip([(19, 13, 11), (70, 5, 9)],
                                        [([1, 3, 2], [3, 1, 2]), ([2, 2, 1], [2, 2, 3])]):
        q = seeded_state(rng, O.EULER3, (mx, my, mz))
        meth = O.method_array(dimensional_split=True)
        dx, dy, dz = 1.0 / mx, 0.9 / my, 1.1 / mz
        dt = 0.2 * min(dx, dy, dz)
        ng = 2
        qbc = O.with_ghosts(q, ng)
        O.bc3(qbc, ng, bcl, bcu)
        qold = qbc.copy(order="F")
        rp = O.make_rp(O.EULER3)
        c1 = O.step3ds(rp, qold, qbc, ng, dx, dy, dz, dt, meth, [4] * 3, 1)
        c2 = O.step3ds(rp, qbc, qbc, ng, dx, dy, dz, dt, meth, [4] * 3, 2)
        c3 = O.step3ds(rp, qbc, qbc, ng, dx, dy, dz, dt, meth, [4] * 3, 3)
        ref = O.interior(qbc, ng)
        qi = q_to_device(q)
        qo = torch.empty_like(qi)
        cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
        K.step3_split(K.make_rp(5, {"gamma": 1.4}), qi, qo, cfl, bcl, bcu, [1, 2, 3], meth,
                      [4] * 3, dx, dy, dz, dt)
        torch.cuda.synchronize()
        out = q_from_device(qo)
        assert np.array_equal(out, ref), ((mx, my, mz), np.abs(out - ref).max())
        assert float(cfl.item()) == max(c1, c2, c3)


# ---------------------------------------------------------------- end to end
def _run_api(rp_module, q, box, bcl, bcu, tfinal, nout=10, problem_data=None, **opts):
    pyclaw = _pyclaw()
    ndim = q.ndim - 1
    lower, upper = box
    dims = [pyclaw.Dimension(lower[k], upper[k], q.shape[k + 1], name="xyz"[k])
            for k in range(ndim)]
    domain = pyclaw.Domain(dims)
    solver_cls = pyclaw.ClawSolver2D if ndim == 2 else pyclaw.ClawSolver3D
    solver = solver_cls(rp_module)
    for k, v in opts.items():
        setattr(solver, k, v)
    for k in range(ndim):
        solver.bc_lower[k] = bcl[k]
        solver.bc_upper[k] = bcu[k]
    state = pyclaw.State(domain, q.shape[0])
    state.problem_data.update(problem_data or {})
    state.q[...] = q
    claw = pyclaw.Controller()
    claw.solution = pyclaw.Solution(state, domain)
    claw.solver = solver
    claw.tfinal = tfinal
    claw.num_output_times = nout
    claw.keep_copy = True
    claw.output_format = None
    claw.run()
    return claw


def _check(claw, run, frame_ref, tol, d):
    final = claw.frames[-1].state.q
    assert claw.solver.status["numsteps"] == run.numsteps
    assert claw.solution.t == run.t
    assert np.array_equal(final, run.q), np.abs(final - run.q).max()
    if frame_ref is not None:
        f0 = final[0] if final.shape[0] > 1 else final
        assert np.prod(d) * np.abs(f0.reshape(-1) - frame_ref).sum() < tol


@pytest.mark.parametrize("key,split,trans", [
    ("classic_dim_split", True, 0), ("classic_unsplit_no_trans", False, 0),
    ("classic_unsplit_trans_inc", False, 1), ("classic_unsplit_trans_cor", False, 2)])
def test_api_advection_fixture(golden_dir, key, split, trans):
    from pyclaw_amd import riemann
    q, d = C.advection_2d()
    cd, cm = (0.45, 0.5) if (not split and trans == 0) else (0.9, 1.0)
    claw = _run_api(riemann.advection_2D, q, ([0, 0], [1, 1]), [2, 2], [2, 2], 2.0,
                    problem_data={"u": 0.5, "v": 1.0}, dimensional_split=split,
                    transverse_waves=trans, limiters=3, cfl_desired=cd, cfl_max=cm)
    run = OracleRun(O.make_rp(O.ADVECTION, u=0.5, v=1.0), q, d, [2, 2], [2, 2],
                    dimensional_split=split, transverse=trans, limiters=3, cfl_desired=cd,
                    cfl_max=cm)
    run.run(2.0, 10)
    _check(claw, run, np.load(f"{golden_dir}/advection_2d_expected.npz")[key], 1e-6, d)


def test_api_euler_quadrants_fixture(golden_dir):
    from pyclaw_amd import riemann
    q, d = C.euler_quadrants()
    claw = _run_api(riemann.euler_4wave_2D, q, ([0, 0], [1, 1]), [1, 1], [1, 1], 0.8,
                    problem_data={"gamma": 1.4}, transverse_waves=2)
    run = OracleRun(O.make_rp(O.EULER, efix=1, gamma=1.4), q, d, [1, 1], [1, 1], limiters=1)
    run.run(0.8, 10)
    _check(claw, run, np.load(f"{golden_dir}/euler_2d_expected.npz")["quadrants_roe"], 1e-6, d)


def test_api_dam_break_fixture(golden_dir):
    from pyclaw_amd import riemann
    q, d = C.radial_dam_break()
    claw = _run_api(riemann.shallow_roe_with_efix_2D, q, ([-2.5, -2.5], [2.5, 2.5]), [1, 1],
                    [3, 3], 2.5,
                    problem_data={"grav": 1.0}, limiters=4, dimensional_split=1)
    run = OracleRun(O.make_rp(O.SHALLOW, efix=1, grav=1.0), q, d, [1, 1], [3, 3], limiters=4)
    run.run(2.5, 10)
    exp = np.load(f"{golden_dir}/shallow_2d_expected.npz")["radialdambreak_classic_roe"]
    _check(claw, run, exp, 1e-5, d)


def test_api_shocktube_3d_fixture(golden_dir):
    from pyclaw_amd import riemann
    q, d = C.shocktube_3d()
    claw = _run_api(riemann.euler_3D, q, ([-1, -1, -1], [1, 1, 1]), [1, 1, 1], [1, 1, 1], 0.2,
                    problem_data={"gamma": 1.4}, limiters=4, cfl_desired=0.8, cfl_max=1.0)
    run = OracleRun(O.make_rp(O.EULER3, efix=1, gamma=1.4), q, d, [1, 1, 1], [1, 1, 1],
                    limiters=4, cfl_desired=0.8, cfl_max=1.0)
    run.run(0.2, 10)
    _check(claw, run, np.load(f"{golden_dir}/euler_3d_expected.npz")["shocktube"], 1e-6, d)


def test_api_unsplit_euler_walls_matches_oracle_and_conserves():
    """Unsplit trans_cor Euler, periodic in x and walls in y, 40 steps: the GPU run
    equals the oracle run bitwise and mass / x-momentum... stay conserved."""
    from pyclaw_amd import riemann
    rng = np.random.default_rng(12)
    q = seeded_state(rng, O.EULER, (90, 70))
    d = [1.0 / 90, 1.0 / 70]
    claw = _run_api(riemann.euler_4wave_2D, q, ([0, 0], [1, 1]), [2, 3], [2, 3], 0.05, nout=2,
                    problem_data={"gamma": 1.4}, dimensional_split=False, transverse_waves=2,
                    limiters=4, dt_initial=0.001)
    run = OracleRun(O.make_rp(O.EULER), q, d, [2, 3], [2, 3], dimensional_split=False,
                    transverse=2, limiters=4, dt_initial=0.001)
    run.run(0.05, 2)
    _check(claw, run, None, None, d)
    final = claw.frames[-1].state.q
    np.testing.assert_allclose(final[0].sum(), q[0].sum(), rtol=1e-13)
    np.testing.assert_allclose(final[3].sum(), q[3].sum(), rtol=1e-13)


def test_api_rejected_step_restores_state():
    """dt_initial far above the CFL limit: the first step is rejected and redone
    (solver.py:641-647) -- same result as the oracle, which copies q_backup."""
    from pyclaw_amd import riemann
    rng = np.random.default_rng(1)
    q = seeded_state(rng, O.SHALLOW, (50, 40))
    d = [1.0 / 50, 1.0 / 40]
    claw = _run_api(riemann.shallow_roe_with_efix_2D, q, ([0, 0], [1, 1]), [1, 1], [1, 1], 0.1,
                    nout=1,
                    problem_data={"grav": 1.0}, dimensional_split=False, transverse_waves=2,
                    limiters=4, dt_initial=0.5)
    run = OracleRun(O.make_rp(O.SHALLOW), q, d, [1, 1], [1, 1], dimensional_split=False,
                    transverse=2, limiters=4, dt_initial=0.5)
    run.run(0.1, 1)
    _check(claw, run, None, None, d)


def test_api_custom_bc_matches_builtin():
    """BC.custom (user function on a host qbc) reproduces the built-in wall."""
    from pyclaw_amd import riemann
    pyclaw = _pyclaw()
    rng = np.random.default_rng(2)
    q = seeded_state(rng, O.EULER, (40, 30))
    box = ([0, 0], [1, 1])

    def wall_lower_x(state, dim, t, qbc, auxbc, num_ghost):
        for i in range(num_ghost):
            qbc[:, i, :] = qbc[:, 2 * num_ghost - 1 - i, :]
            qbc[1, i, :] = -qbc[1, 2 * num_ghost - 1 - i, :]

    a = _run_api(riemann.euler_4wave_2D, q, box, [3, 1], [1, 1], 0.02, nout=1,
                 problem_data={"gamma": 1.4}, dimensional_split=False, transverse_waves=2,
                 limiters=4, dt_initial=0.002)
    b = _run_api(riemann.euler_4wave_2D, q, box, [pyclaw.BC.custom, 1], [1, 1], 0.02, nout=1,
                 problem_data={"gamma": 1.4}, dimensional_split=False, transverse_waves=2,
                 limiters=4, dt_initial=0.002, user_bc_lower=wall_lower_x)
    assert np.array_equal(a.frames[-1].state.q, b.frames[-1].state.q)


@pytest.mark.parametrize("fixed_dt,shape,tfinal", [(False, (80, 60), 0.15),
                                                   (True, (80, 60), 0.15),
                                                   (False, (600, 480), 0.03)])
def test_device_loop_equals_host_loop(fixed_dt, shape, tfinal):
    """The device-controlled loop (clawhip_step2_unsplit_run, graph replays of 16
    attempts plus a remainder) takes the same steps as the host loop: same
    state bit for bit, same t, step count and status, on a grid of a few
    workgroups and on one of a few hundred."""
    from pyclaw_amd import riemann
    print("classic device loop", fixed_dt, shape, flush=True)
    rng = np.random.default_rng(21)
    q = seeded_state(rng, O.EULER, shape)
    if shape[0] > 100:   # milder cell-to-cell noise: the wide grid's run stays finite
        q = 0.8 * q.mean(axis=(1, 2))[:, None, None] + 0.2 * q
    opts = dict(problem_data={"gamma": 1.4}, dimensional_split=False, transverse_waves=2,
                limiters=4, dt_initial=0.02)
    if fixed_dt:
        opts.update(dt_variable=False, dt_initial=0.0005)
    runs = []
    for extra in (dict(device_loop=False), dict(use_graph=False), dict()):
        claw = _run_api(riemann.euler_4wave_2D, q, ([0, 0], [1, 1]), [1, 3], [1, 3], tfinal,
                        nout=3, **opts, **extra)
        runs.append(claw)
    ref = runs[0]
    assert ref.solver.status["numsteps"] >= 20
    for claw in runs[1:]:
        assert claw.solution.t == ref.solution.t
        for key in ("numsteps", "cflmax", "dtmin", "dtmax"):
            assert claw.solver.status[key] == ref.solver.status[key], key
        for k, (fa, fb) in enumerate(zip(claw.frames, ref.frames)):
            assert fa.t == fb.t
            assert np.isfinite(fb.state.q).all(), k
            d = np.abs(fa.state.q - fb.state.q)
            assert np.array_equal(fa.state.q, fb.state.q), (k, d.max(), np.argwhere(d > 0)[:4])


@pytest.mark.parametrize("module", ["acoustics_2D", "acoustics_2D_ptwise"])
def test_api_acoustics_2d_verify_classic(golden_dir, module):
    """examples/acoustics_2d_homogeneous test_classic and test_classic_ptwise
    (split, MC, cfl 0.45 / 0.5, relative 2-norm of the pressure < 1e-3): the
    pointwise module is the same device solver; both runs equal the oracle's."""
    from pyclaw_amd import riemann
    q, d = C.acoustics_2d()
    rp = O.make_rp(O.ACOUSTICS, rho=1.0, bulk=4.0)
    pd = {"rho": 1.0, "bulk": 4.0, "zz": rp.zz, "cc": rp.cc}
    claw = _run_api(getattr(riemann, module), q, ([-1, -1], [1, 1]), [1, 1], [1, 1], 0.12,
                    problem_data=pd, limiters=4, cfl_desired=0.45, cfl_max=0.5,
                    dimensional_split=True, dt_initial=min(d) / rp.cc * 0.45)
    run = OracleRun(rp, q, d, [1, 1], [1, 1], limiters=4, cfl_desired=0.45, cfl_max=0.5,
                    dt_initial=min(d) / rp.cc * 0.45)
    run.run(0.12, 10)
    _check(claw, run, None, None, d)
    exp = np.load(f"{golden_dir}/acoustics_2d_verify_classic.npz")["pressure"]
    final = claw.frames[-1].state.q
    assert np.linalg.norm(exp - final[0]) / np.linalg.norm(exp) < 1e-3
