"""GPU parity of the SharpClaw path (WENO5 component / characteristic, SSP33,
SSP104) against the oracle -- bit for bit in FP64 -- and against the reference's
SharpClaw fixtures."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from oracle.driver import OracleRun
from tests import _cases as C
from tests._cases import seeded_state

pytestmark = pytest.mark.gpu

PD = {O.ACOUSTICS: dict(rho=1.0, bulk=4.0), O.EULER: dict(gamma=1.4),
      O.SHALLOW: dict(grav=1.0), O.ADVECTION: dict(u=0.5, v=-0.7)}


def _orp(kind):
    pd = PD[kind]
    return O.make_rp(kind, efix=1, rho=pd.get("rho", 1.0), bulk=pd.get("bulk", 4.0),
                     gamma=pd.get("gamma", 1.4), grav=pd.get("grav", 1.0),
                     u=pd.get("u", 0.5), v=pd.get("v", 1.0))


@pytest.mark.parametrize("kind,cd", [(O.SHALLOW, 2), (O.SHALLOW, 0), (O.EULER, 2), (O.EULER, 0),
                                     (O.ACOUSTICS, 0), (O.ADVECTION, 0)])
def test_operator_bit_exact(kind, cd):
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device, q_to_device
    rng = np.random.default_rng(7 * kind + cd)
    for (mx, my), (bcl, bcu) in zip([(45, 37), (64, 9), (5, 70), (3, 3)],
                                    [([2, 2], [2, 2]), ([1, 3], [3, 1]), ([3, 2], [1, 2]),
                                     ([1, 1], [3, 3])]):
        if kind == O.ADVECTION:
            bcl, bcu = [2, 1], [2, 1]
        q = seeded_state(rng, kind, (mx, my))
        if kind == O.EULER:
            # cell-to-cell noise of the full admissible range lets WENO5 edge
            # states go negative in pressure (NaN sound speeds); keep it milder
            q = 0.85 * seeded_state(np.random.default_rng(0), O.EULER, (1, 1)).mean(
                axis=(1, 2))[:, None, None] + 0.15 * q
        ng = 3
        qbc = O.with_ghosts(q, ng)
        O.bc2(qbc, ng, bcl, bcu)
        dx, dy = 1.0 / mx, 0.9 / my
        dt = 0.2 * min(dx, dy)
        dq, cref = O.sharp_flux2(_orp(kind), qbc, ng, dx, dy, dt, cd)
        ref = O.interior(dq, ng)
        assert np.isfinite(ref).all() and np.isfinite(cref)
        qi = q_to_device(q)
        out = torch.empty_like(qi)
        cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
        K.sharp2_stage(K.make_rp(kind, PD[kind]), cd, out, cfl, bcl, bcu, [1, 2], dt, dx, dy, qi)
        torch.cuda.synchronize()
        got = q_from_device(out)
        assert np.array_equal(got, ref), ((mx, my), np.abs(got - ref).max())
        assert float(cfl.item()) == cref


def _run_sharp(rp_module, q, box, bcl, bcu, tfinal, nout, pd, **opts):
    import pyclaw_amd as pyclaw
    lower, upper = box
    dims = [pyclaw.Dimension(lower[k], upper[k], q.shape[k + 1], name="xy"[k]) for k in range(2)]
    domain = pyclaw.Domain(dims)
    solver = pyclaw.SharpClawSolver2D(rp_module)
    for k, v in opts.items():
        setattr(solver, k, v)
    for k in range(2):
        solver.bc_lower[k] = bcl[k]
        solver.bc_upper[k] = bcu[k]
    state = pyclaw.State(domain, q.shape[0])
    state.problem_data.update(pd)
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


def test_ssp33_char_weno_swe_matches_oracle():
    from pyclaw_amd import riemann
    rng = np.random.default_rng(21)
    q = seeded_state(rng, O.SHALLOW, (40, 36))
    claw = _run_sharp(riemann.shallow_roe_with_efix_2D, q, ([0, 0], [1, 1]), [2, 2], [2, 2], 0.04,
                      2, {"grav": 1.0}, time_integrator="SSP33", char_decomp=2,
                      cfl_desired=0.5, cfl_max=0.55, dt_initial=0.01)
    run = OracleRun(_orp(O.SHALLOW), q, [1.0 / 40, 1.0 / 36], [2, 2], [2, 2], solver="sharp",
                    time_integrator="SSP33", char_decomp=2, cfl_desired=0.5, cfl_max=0.55,
                    dt_initial=0.01)
    run.run(0.04, 2)
    assert claw.solver.status["numsteps"] == run.numsteps
    assert claw.solution.t == run.t
    got = claw.frames[-1].state.q
    assert np.array_equal(got, run.q), np.abs(got - run.q).max()
    np.testing.assert_allclose(got.sum(axis=(1, 2)), q.sum(axis=(1, 2)), rtol=1e-13, atol=1e-12)


def test_ssp104_advection_fixture(golden_dir):
    from pyclaw_amd import riemann
    q, d = C.advection_2d()
    claw = _run_sharp(riemann.advection_2D, q, ([0, 0], [1, 1]), [2, 2], [2, 2], 2.0, 10,
                      {"u": 0.5, "v": 1.0}, cfl_desired=0.9, cfl_max=1.0)
    run = OracleRun(O.make_rp(O.ADVECTION, u=0.5, v=1.0), q, d, [2, 2], [2, 2], solver="sharp",
                    cfl_desired=0.9, cfl_max=1.0, time_integrator="SSP104")
    run.run(2.0, 10)
    got = claw.frames[-1].state.q
    assert claw.solver.status["numsteps"] == run.numsteps
    assert np.array_equal(got, run.q)
    exp = np.load(f"{golden_dir}/advection_2d_expected.npz")["sharpclaw"]
    assert np.prod(d) * np.abs(got.reshape(-1) - exp).sum() < 1e-6


def test_ssp104_dam_break_fixture(golden_dir):
    from pyclaw_amd import riemann
    q, d = C.radial_dam_break()
    claw = _run_sharp(riemann.shallow_roe_with_efix_2D, q, ([-2.5, -2.5], [2.5, 2.5]), [1, 1],
                      [3, 3], 2.5, 10, {"grav": 1.0})
    got = claw.frames[-1].state.q
    exp = np.load(f"{golden_dir}/shallow_2d_expected.npz")["radialdambreak_sharpclaw_roe"]
    assert np.prod(d) * np.abs(got[0].reshape(-1) - exp).sum() < 1e-5
    run = OracleRun(O.make_rp(O.SHALLOW, efix=1, grav=1.0), q, d, [1, 1], [3, 3], solver="sharp",
                    cfl_desired=2.45, cfl_max=2.5, time_integrator="SSP104")
    run.run(2.5, 10)
    assert claw.solver.status["numsteps"] == run.numsteps
    assert np.array_equal(got, run.q)


def test_fp32_kahan_ssp33_close_to_fp64():
    from pyclaw_amd import riemann
    rng = np.random.default_rng(33)
    q = seeded_state(rng, O.SHALLOW, (64, 64))
    kw = dict(time_integrator="SSP33", char_decomp=2, cfl_desired=0.5, cfl_max=0.55,
              dt_initial=0.004)
    a = _run_sharp(riemann.shallow_roe_with_efix_2D, q, ([0, 0], [1, 1]), [2, 2], [2, 2], 0.05, 1,
                   {"grav": 1.0}, **kw)
    b = _run_sharp(riemann.shallow_roe_with_efix_2D, q, ([0, 0], [1, 1]), [2, 2], [2, 2], 0.05, 1,
                   {"grav": 1.0}, precision="fp32k", **kw)
    qa, qb = a.frames[-1].state.q, b.frames[-1].state.q
    assert np.abs(qa - qb).sum() / np.abs(qa).sum() < 1e-6
    # the Kahan-compensated update keeps the totals at FP64-level drift
    np.testing.assert_allclose(qb.sum(axis=(1, 2)), q.sum(axis=(1, 2)), rtol=1e-6, atol=1e-5)


def test_fp32_kahan_char_weno_flat_regions_stay_finite():
    """Exactly flat regions (h = 1, hu = hv = 0 away from the humps): the FP32
    characteristic weights must not form 0/0 (their (eps + beta)^4 products)."""
    from pyclaw_amd import indicator as I
    out = {}
    for prec in ("fp64", "fp32k"):
        solver = I.swe_solver(prec)
        sol = I.swe_solution(3, 96, 96)
        solver.dt = 0.5 / 96 / 1.6
        solver.setup(sol)
        for _ in range(30):
            solver.evolve_to_time(sol)
        out[prec] = (sol.state.q.copy(), sol.t)
    qa, ta = out["fp64"]
    qb, tb = out["fp32k"]
    assert np.isfinite(qb).all()
    assert abs(tb - ta) < 1e-6 * ta
    assert np.abs(qb - qa).sum() / np.abs(qa).sum() < 1e-5


def test_ssplmm_k2_acoustics_fixture(golden_dir):
    """examples/acoustics_2d_homogeneous test_sharpclaw_multistep through the
    PyClaw API: SSPLMMk2 with lmm_steps 3 (SSP22 starting steps, then the
    multistep update), equal to the oracle run bit for bit and to the reference's
    verify_sharpclaw_lmm.txt to round-off (bar: relative 1e-3)."""
    from pyclaw_amd import riemann
    q, d = C.acoustics_2d()
    pd = {"rho": 1.0, "bulk": 4.0, "cc": 2.0, "zz": 2.0}
    claw = _run_sharp(riemann.acoustics_2D, q, ([-1, -1], [1, 1]), [1, 1], [1, 1], 0.12, 10, pd,
                      time_integrator="SSPLMMk2", lmm_steps=3, cfl_max=0.25, cfl_desired=0.24,
                      dt_initial=min(d) / 2.0 * 0.24)
    rp = O.make_rp(O.ACOUSTICS, rho=1.0, bulk=4.0)
    run = OracleRun(rp, q, d, [1, 1], [1, 1], solver="sharp", cfl_desired=0.24, cfl_max=0.25,
                    time_integrator="SSPLMMk2", lmm_steps=3, dt_initial=min(d) / rp.cc * 0.24)
    run.run(0.12, 10)
    got = claw.frames[-1].state.q
    assert claw.solver.status["numsteps"] == run.numsteps
    assert np.array_equal(got, run.q), np.abs(got - run.q).max()
    exp = np.load(f"{golden_dir}/acoustics_2d_verify_sharpclaw.npz")["sharpclaw_lmm"]
    rel = np.linalg.norm(exp - got[0]) / np.linalg.norm(exp)
    print("SSPLMMk2 relative error", rel)
    assert rel < 1e-12


@pytest.mark.parametrize("ti,steps", [("SSPLMMk3", 4), ("SSPLMMk3", 5), ("SSPLMMk2", 6),
                                      ("SSP22", 0), ("Euler", 0)])
def test_more_integrators_match_oracle(ti, steps):
    """SSPLMMk3 (4 and 5 steps), SSPLMMk2 with more steps (their default CFL
    limits), SSP22 and forward Euler: whole runs on shallow water equal the
    oracle's bit for bit (same accepted steps, same state)."""
    from pyclaw_amd import riemann
    rng = np.random.default_rng(55)
    q = seeded_state(rng, O.SHALLOW, (40, 30))
    opts = dict(time_integrator=ti, dt_initial=0.0005)
    kw = dict(solver="sharp", time_integrator=ti, dt_initial=0.0005)
    if steps:
        opts["lmm_steps"] = kw["lmm_steps"] = steps
        sc = (steps - 2.) / (steps - 1.) if ti == "SSPLMMk2" else (steps - 3.) / (steps - 1.)
        cd, cm = (0.14 * sc, 0.15 * sc) if ti == "SSPLMMk2" else (0.48 * sc, 0.5 * sc)
    else:
        cd, cm = (0.45, 0.5) if ti == "SSP22" else (0.2, 0.25)
        opts.update(cfl_desired=cd, cfl_max=cm)
    claw = _run_sharp(riemann.shallow_roe_with_efix_2D, q, ([0, 0], [1, 1]), [2, 1], [2, 3],
                      0.03, 2, {"grav": 1.0}, **opts)
    run = OracleRun(O.make_rp(O.SHALLOW, efix=1, grav=1.0), q, [1.0 / 40, 1.0 / 30], [2, 1],
                    [2, 3], cfl_desired=cd, cfl_max=cm, **kw)
    run.run(0.03, 2)
    got = claw.frames[-1].state.q
    assert claw.solver.status["numsteps"] == run.numsteps
    assert np.array_equal(got, run.q), np.abs(got - run.q).max()


RK3_A = [[0., 0., 0.], [0.5, 0., 0.], [-1., 2., 0.]]   # Kutta's third-order method
RK3_B = [1. / 6., 2. / 3., 1. / 6.]


@pytest.mark.parametrize("ti", ["RK", "LMM"])
def test_general_integrators_match_oracle(ti):
    """The general explicit Runge-Kutta method (solver.a, b: Kutta's third-order
    tableau) with variable dt, and the general linear multistep method
    (solver.alpha, beta: the 3-step Adams-Bashforth weights, SSP104 starting
    steps, fixed dt): whole runs on shallow water equal the oracle's bit for bit."""
    from pyclaw_amd import riemann
    rng = np.random.default_rng(56)
    q = seeded_state(rng, O.SHALLOW, (40, 30))
    if ti == "RK":
        opts = dict(time_integrator="RK", a=RK3_A, b=RK3_B, c=[0., 0.5, 1.], cfl_desired=0.45,
                    cfl_max=0.5, dt_initial=0.0005)
        kw = dict(time_integrator="RK", rk_a=RK3_A, rk_b=RK3_B, rk_c=[0., 0.5, 1.],
                  cfl_desired=0.45, cfl_max=0.5, dt_initial=0.0005)
    else:
        al, be = [0., 0., 1.], [5. / 12., -16. / 12., 23. / 12.]
        opts = dict(time_integrator="LMM", alpha=al, beta=be, lmm_steps=3, dt_variable=False,
                    dt_initial=0.0005, cfl_max=0.5, cfl_desired=0.45)
        kw = dict(time_integrator="LMM", lmm_alpha=al, lmm_beta=be, lmm_steps=3,
                  dt_variable=False, dt_initial=0.0005, cfl_max=0.5, cfl_desired=0.45)
    claw = _run_sharp(riemann.shallow_roe_with_efix_2D, q, ([0, 0], [1, 1]), [2, 1], [2, 3],
                      0.01, 2, {"grav": 1.0}, **opts)
    run = OracleRun(O.make_rp(O.SHALLOW, efix=1, grav=1.0), q, [1.0 / 40, 1.0 / 30], [2, 1],
                    [2, 3], solver="sharp", **kw)
    run.run(0.01, 2)
    got = claw.frames[-1].state.q
    assert claw.solver.status["numsteps"] == run.numsteps
    assert np.array_equal(got, run.q), np.abs(got - run.q).max()


@pytest.mark.parametrize("kind,meqn,maux", [(O.VC_ACOUSTICS, 3, 2), (O.VC_ACOUSTICS, 3, 3),
                                            (O.VC_ADVECTION_2D, 1, 2), (O.VC_ADVECTION, 2, 2)])
def test_operator_aux_bit_exact(kind, meqn, maux):
    """The SharpClaw operator with aux fields (interface solves with both cells'
    aux, in-cell solves with the cell's own) equals the oracle's, whose aux
    operator equals the reference's flux2 (tests/test_oracle.py); compact aux
    with boundary codes and aux with stored ghosts; FP32 close."""
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device, q_to_device
    rng = np.random.default_rng(61 + kind + maux)
    lo, hi = (0.5, 2.0) if kind == O.VC_ACOUSTICS else (-1.5, 1.5)
    for (mx, my), (bcl, bcu) in zip([(45, 37), (64, 9)], [([3, 1], [2, 3]), ([1, 2], [1, 2])]):
        print("sharp aux", kind, meqn, maux, mx, my, flush=True)
        q = np.asfortranarray(rng.uniform(-1, 1, (meqn, mx, my)))
        aux = np.asfortranarray(rng.uniform(lo, hi, (maux, mx, my)))
        ng = 3
        qbc, abc = O.with_ghosts(q, ng), O.with_ghosts(aux, ng)
        O.bc2(qbc, ng, bcl, bcu)
        O.bc2(abc, ng, bcl, bcu, reflect=(-1, -1))
        dx, dy = 1.0 / mx, 0.9 / my
        dt = 0.1 * min(dx, dy)
        rp_o = O.make_rp(kind, meqn=meqn)
        dq, cref = O.sharp_flux2(rp_o, qbc, ng, dx, dy, dt, 0, aux=abc)
        ref = O.interior(dq, ng)
        rp = K.make_rp(kind, {}, meqn=meqn)
        for given in (False, True):
            out = torch.empty_like(q_to_device(q))
            cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
            K.sharp2_stage(rp, 0, out, cfl, bcl, bcu, [1, 2], dt, dx, dy, q_to_device(q),
                           qaux=q_to_device(abc if given else aux),
                           qaux_bc_lower=[4, 4] if given else bcl,
                           qaux_bc_upper=[4, 4] if given else bcu)
            torch.cuda.synchronize()
            got = q_from_device(out)
            assert np.array_equal(got, ref), (given, np.abs(got - ref).max())
            assert float(cfl.item()) == cref
        out32 = torch.empty_like(q_to_device(q, 1))
        cfl32 = torch.zeros(1, dtype=torch.float64, device="cuda")
        K.sharp2_stage(rp, 0, out32, cfl32, bcl, bcu, [1, 2], dt, dx, dy, q_to_device(q, 1),
                       qaux=q_to_device(aux, 1), qaux_bc_lower=bcl, qaux_bc_upper=bcu)
        torch.cuda.synchronize()
        assert np.abs(q_from_device(out32) - ref).max() <= 1e-5 * max(np.abs(ref).max(), 1e-30)


@pytest.mark.parametrize("ti,cd,cm", [("SSP104", 2.45, 2.5), ("SSPLMMk2", 0.24, 0.25)])
def test_acoustics_2d_variable_sharpclaw_fixture(golden_dir, ti, cd, cm):
    """examples/acoustics_2d_variable test_sharpclaw / test_sharpclaw_multistep
    through the PyClaw API (vc_acoustics_2D with aux density and sound speed,
    walls / extrapolation for q and aux): equal to the oracle run bit for bit,
    and within the fixture's 0.1 of pressure_sharpclaw.txt."""
    import pyclaw_amd as pyclaw
    from pyclaw_amd import riemann
    q0, aux0, d = C.acoustics_2d_variable()
    solver = pyclaw.SharpClawSolver2D(riemann.vc_acoustics_2D)
    solver.time_integrator = ti
    if ti == "SSPLMMk2":
        solver.lmm_steps = 3
        solver.cfl_max = cm
        solver.cfl_desired = cd
    solver.bc_lower[0] = pyclaw.BC.wall
    solver.bc_upper[0] = pyclaw.BC.extrap
    solver.bc_lower[1] = pyclaw.BC.wall
    solver.bc_upper[1] = pyclaw.BC.extrap
    solver.aux_bc_lower[0] = pyclaw.BC.wall
    solver.aux_bc_upper[0] = pyclaw.BC.extrap
    solver.aux_bc_lower[1] = pyclaw.BC.wall
    solver.aux_bc_upper[1] = pyclaw.BC.extrap
    domain = pyclaw.Domain([pyclaw.Dimension(-1.0, 1.0, 50, name="x"),
                            pyclaw.Dimension(-1.0, 1.0, 50, name="y")])
    state = pyclaw.State(domain, 3, 2)
    state.aux[...] = aux0
    state.q[...] = q0
    claw = pyclaw.Controller()
    claw.keep_copy = True
    claw.output_format = None
    claw.solution = pyclaw.Solution(state, domain)
    claw.solver = solver
    claw.tfinal = 0.6
    claw.num_output_times = 20
    claw.run()
    run = OracleRun(O.make_rp(O.VC_ACOUSTICS), q0, d, [3, 3], [1, 1], solver="sharp",
                    time_integrator=ti, cfl_desired=cd, cfl_max=cm, lmm_steps=3, aux=aux0,
                    aux_bc_lower=[3, 3], aux_bc_upper=[1, 1])
    run.run(0.6, 20)
    final = claw.frames[-1].state.q
    assert claw.solver.status["numsteps"] == run.numsteps
    assert np.array_equal(final, run.q), np.abs(final - run.q).max()
    p = np.load(f"{golden_dir}/acoustics_2d_variable_pressure_sharpclaw.npz")["pressure"]
    assert np.abs(p.reshape(-1) - final[0].reshape(-1)).max() < 1e-1


def _timed_callback_runs(ti, precision="fp64"):
    """Shallow water with a time-dependent inflow at x lower (custom BC) and a
    time-dependent source (dq_src): through the API and in the oracle, where
    both callbacks read the time of the stage's state."""
    import pyclaw_amd as pyclaw
    from pyclaw_amd import riemann
    rng = np.random.default_rng(91)
    q = seeded_state(rng, O.SHALLOW, (30, 24))
    S = np.array([0.2, -0.1, 0.05])[:, None, None]

    def inflow(t):
        return np.array([1.2 + 0.5 * t, 0.3 - t, 0.1 * t])

    def api_bc(state, dim, t, qbc, auxbc, num_ghost):
        for i in range(num_ghost):
            qbc[:, i, :] = inflow(t)[:, None]

    def api_src(solver, state, dt):
        return dt * (state.t * S - 0.1 * state.q)

    def orc_bc(run, dim, qbc):
        for i in range(run.ng):
            qbc[:, i, :] = inflow(run.stage_t)[:, None]

    def orc_src(run, qq, dt):
        return dt * (run.stage_t * S - 0.1 * qq)

    kw = dict(time_integrator=ti, cfl_desired=0.45 if ti != "SSP104" else 2.4,
              cfl_max=0.5 if ti != "SSP104" else 2.5)
    claw = _run_sharp(riemann.shallow_roe_with_efix_2D, q, ([0, 0], [1, 1]),
                      [pyclaw.BC.custom, 1], [1, 1], 0.3, 2, {"grav": 1.0},
                      user_bc_lower=api_bc, dq_src=api_src, dt_initial=0.005,
                      precision=precision, **kw)
    run = OracleRun(_orp(O.SHALLOW), q, [1.0 / 30, 1.0 / 24], [O.CUSTOM, 1], [1, 1],
                    solver="sharp", dt_initial=0.005, user_bc_lower=orc_bc, dq_src=orc_src,
                    **kw)
    run.run(0.3, 2)
    return claw, run


@pytest.mark.parametrize("ti", ["SSP33", "SSP104", "SSP22"])
def test_time_dependent_bc_and_source_match_oracle(ti):
    """Custom BC functions and dq_src see each stage's own time (the registers'
    t, sharpclaw/solver.py:313-338, 391, 409-427): whole FP64 runs with a
    time-dependent inflow and source equal the oracle's bit for bit (the
    oracle's stage times are pinned by quadrature in tests/test_oracle.py)."""
    claw, run = _timed_callback_runs(ti)
    got = claw.frames[-1].state.q
    assert claw.solver.status["numsteps"] == run.numsteps
    assert np.array_equal(got, run.q), np.abs(got - run.q).max()


def test_fp32_kahan_with_host_callbacks_close_to_fp64():
    """FP32 + Kahan SSP33 whose stage inputs are formed on the host (custom BC and
    dq_src): its second stage still takes the step's dt*dq_dt.  After ~27 steps
    the FP32 run is within 5e-5 relative L1 and 1e-4 relative max of the FP64
    run (which equals the oracle's); dropping dt*dq_dt from the second stage
    moves the state by ~0.25*dt*|dq/dt|, above 1e-3 relative."""
    a, run = _timed_callback_runs("SSP33")
    b, _ = _timed_callback_runs("SSP33", precision="fp32k")
    qa, qb = a.frames[-1].state.q, b.frames[-1].state.q
    assert np.array_equal(qa, run.q)
    assert abs(b.solution.t - a.solution.t) < 1e-6 * a.solution.t
    assert np.abs(qa - qb).sum() / np.abs(qa).sum() < 5e-5
    assert np.abs(qa - qb).max() / np.abs(qa).max() < 1e-4
