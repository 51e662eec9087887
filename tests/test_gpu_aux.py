"""Aux fields on the device: the dimension-split 2D step with the variable-
coefficient advection solver of examples/advection_reaction_2d (edge velocities
in two aux fields), and the unsplit step with aux fields and a capacity field
(vc_advection_2D on the mapped annulus of examples/advection_2d_annulus), each
against the oracle (whose aux steps equal the reference's own Fortran step2ds /
step2 with that example's rpn2, tests/test_oracle.py), through the PyClaw API
(the examples' fixtures), and through the fmod drop-in."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from oracle.driver import OracleRun
from tests import _cases as C

pytestmark = pytest.mark.gpu


def _dev(q, dtype=0):
    from pyclaw_amd.device import q_to_device
    return q_to_device(q, dtype)


@pytest.mark.parametrize("meqn,bcl,bcu", [
    (2, [1, 3], [2, 1]), (1, [3, 2], [3, 2]), (3, [2, 1], [2, 3]), (4, [1, 1], [1, 1])])
def test_split_step_aux_bit_exact(meqn, bcl, bcu):
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device
    rng = np.random.default_rng(30 + meqn)
    mx, my, ng = 70, 45, 2
    q = np.asfortranarray(rng.uniform(-1.0, 2.0, (meqn, mx, my)))
    aux = np.asfortranarray(rng.uniform(-1.5, 1.5, (2, mx, my)))
    dx, dy, dt = 1.0 / mx, 1.0 / my, 0.004
    meth = O.method_array(dimensional_split=True, num_aux=2)
    # oracle: BC fill of q and aux (aux never negated), sweeps 1 and 2
    qbc = O.with_ghosts(q, ng)
    abc = O.with_ghosts(aux, ng)
    O.bc2(qbc, ng, bcl, bcu)
    O.bc2(abc, ng, bcl, bcu, reflect=(-1, -1))
    rp_o = O.make_rp(O.VC_ADVECTION, meqn=meqn)
    qold = qbc.copy(order="F")
    c1 = O.step2(rp_o, qold, qbc, ng, dx, dy, dt, meth, [4], ids=1, aux=abc)
    c2 = O.step2(rp_o, qbc, qbc, ng, dx, dy, dt, meth, [4], ids=2, aux=abc)
    ref = O.interior(qbc, ng)
    rp = K.make_rp(6, {}, meqn=meqn)
    out = torch.empty_like(_dev(q))
    cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
    K.step2_split(rp, _dev(q), out, cfl, bcl, bcu, [1, 2], meth, [4], dx, dy, dt,
                  aux=_dev(aux), aux_bc_lower=bcl, aux_bc_upper=bcu)
    torch.cuda.synchronize()
    got = q_from_device(out)
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    assert float(cfl.item()) == max(c1, c2)
    # FP32 + Kahan: close to FP64
    q32 = _dev(q, 1)
    out32 = torch.empty_like(q32)
    comp = torch.zeros_like(q32)
    comp_out = torch.empty_like(q32)
    cfl32 = torch.zeros(1, dtype=torch.float64, device="cuda")
    K.step2_split(rp, q32, out32, cfl32, bcl, bcu, [1, 2], meth, [4], dx, dy, dt,
                  comp_in=comp, comp_out=comp_out, aux=_dev(aux, 1), aux_bc_lower=bcl,
                  aux_bc_upper=bcu)
    torch.cuda.synchronize()
    got32 = q_from_device(out32)
    assert np.abs(got32 - ref).max() <= 1e-5 * np.abs(ref).max()


def test_split_step_aux_required():
    """A solver with aux fields needs them; the vector solver has no transverse
    split, so the unsplit step takes it only with transverse_waves = 0."""
    from pyclaw_amd import kernels as K
    from pyclaw_amd._lib import ClawhipError
    q = _dev(np.ones((2, 16, 16), order="F"))
    out = torch.empty_like(q)
    cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
    rp = K.make_rp(6, {}, meqn=2)
    meth = O.method_array(dimensional_split=True, num_aux=2)
    with pytest.raises(ClawhipError):
        K.step2_split(rp, q, out, cfl, [1, 1], [1, 1], [1, 2], meth, [4], 0.1, 0.1, 0.01)
    with pytest.raises(ClawhipError):
        K.step2_unsplit(rp, q, out, cfl, [1, 1], [1, 1], [1, 2], O.method_array(num_aux=2),
                        [4], 0.1, 0.1, 0.01)


def test_api_advection_reaction_fixture(golden_dir):
    """examples/advection_reaction_2d (test_advection_reaction.py): ClawSolver2D with
    the example's Riemann solver, aux edge velocities reset by before_step every
    step, the reaction by step_source; bitwise equal to the oracle run and within
    the fixture's check_diff tolerance (reltol 1e-5, 2-norm)."""
    import pyclaw_amd as pyclaw
    from pyclaw_amd import riemann
    q0, aux0, d, set_vel, source = C.advection_reaction_2d()
    solver = pyclaw.ClawSolver2D(riemann.vc_advection_vector_2D)
    solver.dimensional_split = 1
    solver.all_bcs = pyclaw.BC.extrap
    for k in range(2):
        solver.aux_bc_lower[k] = pyclaw.BC.extrap
        solver.aux_bc_upper[k] = pyclaw.BC.extrap
    solver.num_eqn = 2
    solver.num_waves = 1
    domain = pyclaw.Domain((0., 0.), (1., 1.), (100, 100))
    state = pyclaw.State(domain, 2, 2)
    state.aux[...] = aux0

    def before_step(solver, state):
        state.aux[...] = set_vel(state.t, solver.dt)

    def step_source(solver, state, dt):
        source(state.q, dt)

    solver.before_step = before_step
    solver.step_source = step_source
    solver.source_split = 1
    state.q[...] = q0
    claw = pyclaw.Controller()
    claw.tfinal = 4.0
    claw.solution = pyclaw.Solution(state, domain)
    claw.solver = solver
    claw.keep_copy = True
    claw.output_format = None
    claw.run()

    def before(run):
        run.aux = set_vel(run.t, run.dt)

    def src(run, dt):
        source(run.q, dt)

    run = OracleRun(O.make_rp(O.VC_ADVECTION, meqn=2), q0, d, [1, 1], [1, 1],
                    dimensional_split=True, limiters=1, aux=aux0, aux_bc_lower=[1, 1],
                    aux_bc_upper=[1, 1], before_step=before, step_source=src)
    run.run(4.0, 10)
    final = claw.frames[-1].state.q
    assert claw.solver.status["numsteps"] == run.numsteps
    assert claw.solution.t == run.t
    assert np.array_equal(final, run.q), np.abs(final - run.q).max()
    exp = np.load(f"{golden_dir}/advection_reaction_2d_expected.npz")["density"]
    assert np.linalg.norm(exp - final[0]) / np.linalg.norm(exp) < 1e-5


def test_fmod_step2ds_aux():
    """pyclaw_amd.fmod.classic2.step2ds with aux (the f2py calling convention):
    both sweeps equal the oracle's."""
    from pyclaw_amd.fmod import classic2
    from pyclaw_amd import riemann
    rng = np.random.default_rng(9)
    meqn, mx, my, ng = 2, 33, 21, 2
    qbc = O.with_ghosts(np.asfortranarray(rng.uniform(0, 1, (meqn, mx, my))), ng)
    abc = O.with_ghosts(np.asfortranarray(rng.uniform(-1, 1, (2, mx, my))), ng)
    O.bc2(qbc, ng, [1, 2], [1, 2])
    O.bc2(abc, ng, [1, 2], [1, 2], reflect=(-1, -1))
    meth = np.array(O.method_array(dimensional_split=True, num_aux=2), dtype=np.int32)
    dx, dy, dt = 1.0 / mx, 1.0 / my, 0.01
    rp_o = O.make_rp(O.VC_ADVECTION, meqn=meqn)
    for ids in (1, 2):
        ref = qbc.copy(order="F")
        c_ref = O.step2(rp_o, qbc.copy(order="F"), ref, ng, dx, dy, dt, list(meth), [4],
                        ids=ids, aux=abc)
        qnew = qbc.copy(order="F")
        out, cfl = classic2.step2ds(max(mx, my), ng, mx, my, qbc.copy(order="F"), qnew, abc,
                                    dx, dy, dt, meth, np.array([4]), None, None, None, None,
                                    ids, False, riemann.vc_advection_vector_2D.rpn2._cpointer,
                                    riemann.vc_advection_vector_2D.rpt2._cpointer)
        sl = (slice(None), slice(ng, -ng), slice(ng, -ng))
        assert np.array_equal(out[sl], ref[sl]) and cfl == c_ref


_UNSPLIT_CASES = [
    # kind, meqn, order, trans, limiter, index_capa, num_aux, bc_lower, bc_upper
    (7, 1, 2, 2, 3, 2, 3, [1, 2], [1, 2]),
    (7, 1, 2, 1, 4, -1, 2, [3, 1], [2, 3]),
    (7, 1, 1, 2, 1, 2, 3, [2, 3], [2, 1]),
    (7, 1, 2, 0, 2, 0, 3, [1, 1], [3, 3]),
    (6, 2, 2, 0, 4, 2, 3, [1, 1], [3, 2]),
    (6, 4, 2, 0, 3, -1, 2, [2, 2], [2, 2]),
    (8, 3, 2, 1, 4, -1, 2, [3, 3], [1, 1]),
    (8, 3, 2, 2, 1, 2, 3, [3, 2], [1, 2]),
    (8, 3, 1, 1, 4, -1, 2, [2, 1], [2, 3]),
]


@pytest.mark.parametrize("case", _UNSPLIT_CASES)
def test_unsplit_step_aux_capa_bit_exact(case):
    """One unsplit step with aux fields (and a capacity field) equals the
    oracle's step2 bit for bit, CFL included; the FP32 + Kahan step is close."""
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device
    kind, meqn, order, trans, lim, capa, maux, bcl, bcu = case
    rng = np.random.default_rng(40 + kind + meqn + order + trans)
    mx, my, ng = 77, 50, 2
    q = np.asfortranarray(rng.uniform(-1.0, 2.0, (meqn, mx, my)))
    aux = rng.uniform(-1.5, 1.5, (maux, mx, my))
    if kind == 8:      # density and sound speed
        aux[:2] = rng.uniform(0.5, 2.0, (2, mx, my))
    if capa >= 0:
        aux[capa] = rng.uniform(0.5, 2.0, (mx, my))
    aux = np.asfortranarray(aux)
    dx, dy, dt = 1.0 / mx, 1.0 / my, 0.002
    meth = O.method_array(order=order, transverse=trans, num_aux=maux, index_capa=capa)
    qbc = O.with_ghosts(q, ng)
    abc = O.with_ghosts(aux, ng)
    O.bc2(qbc, ng, bcl, bcu)
    O.bc2(abc, ng, bcl, bcu, reflect=(-1, -1))
    qold = qbc.copy(order="F")
    cref = O.step2(O.make_rp(kind, meqn=meqn), qold, qbc, ng, dx, dy, dt, meth, [lim], aux=abc)
    ref = O.interior(qbc, ng)
    rp = K.make_rp(kind, {}, meqn=meqn)
    print("unsplit aux case", case, flush=True)
    for given in (False, True):   # aux compact with BC codes, or with stored ghosts
        out = torch.empty_like(_dev(q))
        cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
        a_dev = _dev(abc if given else aux)
        K.step2_unsplit(rp, _dev(q), out, cfl, bcl, bcu, [1, 2], meth, [lim], dx, dy, dt,
                        aux=a_dev, aux_bc_lower=[4, 4] if given else bcl,
                        aux_bc_upper=[4, 4] if given else bcu)
        torch.cuda.synchronize()
        got = q_from_device(out)
        assert np.array_equal(got, ref), (given, np.abs(got - ref).max())
        assert float(cfl.item()) == cref
    q32 = _dev(q, 1)
    out32 = torch.empty_like(q32)
    comp = torch.zeros_like(q32)
    comp_out = torch.empty_like(q32)
    cfl32 = torch.zeros(1, dtype=torch.float64, device="cuda")
    K.step2_unsplit(rp, q32, out32, cfl32, bcl, bcu, [1, 2], meth, [lim], dx, dy, dt,
                    comp_in=comp, comp_out=comp_out, aux=_dev(aux, 1), aux_bc_lower=bcl,
                    aux_bc_upper=bcu)
    torch.cuda.synchronize()
    assert np.abs(q_from_device(out32) - ref).max() <= 1e-5 * np.abs(ref).max()
    assert abs(float(cfl32.item()) - cref) <= 1e-5 * cref


def test_unsplit_aux_argument_checks():
    """The unsplit entry point refuses an aux count outside [2, 3], a capacity
    index beyond the aux fields, and aux for a solver without aux fields."""
    from pyclaw_amd import kernels as K
    from pyclaw_amd._lib import ClawhipError
    q = _dev(np.ones((1, 16, 16), order="F"))
    out = torch.empty_like(q)
    cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
    rp = K.make_rp(7, {})
    aux4 = _dev(np.ones((4, 16, 16), order="F"))
    aux3 = _dev(np.ones((3, 16, 16), order="F"))
    with pytest.raises(ClawhipError):
        K.step2_unsplit(rp, q, out, cfl, [1, 1], [1, 1], [1, 2], O.method_array(num_aux=4),
                        [3], 0.1, 0.1, 0.01, aux=aux4, aux_bc_lower=[1, 1], aux_bc_upper=[1, 1])
    with pytest.raises(ClawhipError):
        K.step2_unsplit(rp, q, out, cfl, [1, 1], [1, 1], [1, 2],
                        O.method_array(num_aux=3)[:5] + [4, 3], [3], 0.1, 0.1, 0.01, aux=aux3,
                        aux_bc_lower=[1, 1], aux_bc_upper=[1, 1])
    with pytest.raises(ClawhipError):
        K.step2_unsplit(K.make_rp(4, {}), q, out, cfl, [1, 1], [1, 1], [1, 2],
                        O.method_array(num_aux=3), [3], 0.1, 0.1, 0.01, aux=aux3,
                        aux_bc_lower=[1, 1], aux_bc_upper=[1, 1])
    torch.cuda.synchronize()


def _annulus_claw(c, precision="fp64"):
    """examples/advection_2d_annulus/advection_annulus.py:150-213 through the
    PyClaw API (mapc2p, custom aux ghosts, index_capa 2)."""
    import pyclaw_amd as pyclaw
    from pyclaw_amd import riemann
    solver = pyclaw.ClawSolver2D(riemann.vc_advection_2D)
    solver.dimensional_split = False
    solver.transverse_waves = 2
    solver.order = 2
    solver.precision = precision
    solver.bc_lower[0] = pyclaw.BC.extrap
    solver.bc_upper[0] = pyclaw.BC.extrap
    solver.bc_lower[1] = pyclaw.BC.periodic
    solver.bc_upper[1] = pyclaw.BC.periodic
    solver.aux_bc_lower[0] = pyclaw.BC.custom
    solver.aux_bc_upper[0] = pyclaw.BC.custom
    solver.user_aux_bc_lower = c["ghost_lower"]
    solver.user_aux_bc_upper = c["ghost_upper"]
    solver.aux_bc_lower[1] = pyclaw.BC.periodic
    solver.aux_bc_upper[1] = pyclaw.BC.periodic
    solver.dt_initial = 0.1
    solver.cfl_max = 0.5
    solver.cfl_desired = 0.4
    solver.limiters = pyclaw.limiters.tvd.vanleer
    r = pyclaw.Dimension(c["lower"][0], c["upper"][0], c["num_cells"][0], name="r")
    theta = pyclaw.Dimension(c["lower"][1], c["upper"][1], c["num_cells"][1], name="theta")
    domain = pyclaw.Domain([r, theta])
    domain.grid.mapc2p = c["mapc2p"]
    domain.grid.num_ghost = solver.num_ghost
    state = pyclaw.State(domain, 1)
    X, Y = state.grid.p_centers
    state.q[0, :, :] = c["q"][0]
    dx, dy = state.grid.delta
    Xn, Yn = state.grid.p_nodes
    state.aux = c["edge_aux"](Xn, Yn, dx, dy)
    state.index_capa = 2
    claw = pyclaw.Controller()
    claw.tfinal = 1.0
    claw.solution = pyclaw.Solution(state, domain)
    claw.solver = solver
    claw.output_format = None
    claw.keep_copy = True
    return claw


def test_api_advection_annulus_fixture(golden_dir):
    """examples/advection_2d_annulus (test_2d_advection_annulus.py: area-weighted
    L1 < 1e-6 against expected_sols 'classic'): bitwise equal to the oracle run,
    which itself meets the fixture; the FP32 + Kahan run also meets it."""
    from tests.test_oracle import annulus_l1, annulus_run
    c = C.advection_annulus()
    claw = _annulus_claw(c)
    np.testing.assert_array_equal(claw.solution.state.aux, c["aux"])
    claw.run()
    run = annulus_run(c)
    run.run(1.0, 10)
    final = claw.frames[claw.num_output_times].state.get_q_global()
    assert claw.solver.status["numsteps"] == run.numsteps
    assert claw.solution.t == run.t
    assert np.array_equal(final, run.q), np.abs(final - run.q).max()
    err = annulus_l1(c, final, golden_dir)
    print("annulus L1 fp64", err)
    assert err < 1e-6
    claw32 = _annulus_claw(c, "fp32k")
    claw32.run()
    err32 = annulus_l1(c, claw32.frames[claw32.num_output_times].state.get_q_global(), golden_dir)
    print("annulus L1 fp32k", err32)
    assert err32 < 1e-4


def test_fmod_step2_aux_capa():
    """pyclaw_amd.fmod.classic2.step2 with aux and index_capa (the f2py calling
    convention, ghosts filled by the caller): equal to the oracle's step2."""
    from pyclaw_amd.fmod import classic2
    from pyclaw_amd import riemann
    rng = np.random.default_rng(19)
    meqn, mx, my, ng = 1, 29, 23, 2
    qbc = O.with_ghosts(np.asfortranarray(rng.uniform(0, 1, (meqn, mx, my))), ng)
    aux = rng.uniform(-1, 1, (3, mx, my))
    aux[2] = rng.uniform(0.5, 2.0, (mx, my))
    abc = O.with_ghosts(np.asfortranarray(aux), ng)
    O.bc2(qbc, ng, [1, 2], [1, 2])
    O.bc2(abc, ng, [1, 2], [1, 2], reflect=(-1, -1))
    meth = np.array(O.method_array(num_aux=3, index_capa=2), dtype=np.int32)
    dx, dy, dt = 1.0 / mx, 1.0 / my, 0.01
    ref = qbc.copy(order="F")
    c_ref = O.step2(O.make_rp(O.VC_ADVECTION_2D), qbc.copy(order="F"), ref, ng, dx, dy, dt,
                    list(meth), [4], aux=abc)
    out, cfl = classic2.step2(max(mx, my), ng, mx, my, qbc.copy(order="F"), qbc.copy(order="F"),
                              abc, dx, dy, dt, meth, np.array([4]), None, None, None, None,
                              False, riemann.vc_advection_2D.rpn2._cpointer,
                              riemann.vc_advection_2D.rpt2._cpointer)
    sl = (slice(None), slice(ng, -ng), slice(ng, -ng))
    assert np.array_equal(out[sl], ref[sl]) and cfl == c_ref


@pytest.mark.parametrize("bcl,bcu", [([3, 3], [1, 1]), ([2, 1], [2, 3])])
def test_split_step_vc_acoustics_bit_exact(bcl, bcu):
    """vc_acoustics_2D in the dimension-split step: both sweeps equal the
    oracle's bit for bit, CFL included."""
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device
    rng = np.random.default_rng(77)
    mx, my, ng = 66, 41, 2
    q = np.asfortranarray(rng.uniform(-1.0, 1.0, (3, mx, my)))
    aux = np.asfortranarray(rng.uniform(0.5, 2.0, (2, mx, my)))
    dx, dy, dt = 1.0 / mx, 1.0 / my, 0.003
    meth = O.method_array(dimensional_split=True, num_aux=2)
    qbc, abc = O.with_ghosts(q, ng), O.with_ghosts(aux, ng)
    O.bc2(qbc, ng, bcl, bcu)
    O.bc2(abc, ng, bcl, bcu, reflect=(-1, -1))
    rp_o = O.make_rp(O.VC_ACOUSTICS)
    qold = qbc.copy(order="F")
    c1 = O.step2(rp_o, qold, qbc, ng, dx, dy, dt, meth, [4, 4], ids=1, aux=abc)
    c2 = O.step2(rp_o, qbc, qbc, ng, dx, dy, dt, meth, [4, 4], ids=2, aux=abc)
    ref = O.interior(qbc, ng)
    out = torch.empty_like(_dev(q))
    cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
    K.step2_split(K.make_rp(8, {}), _dev(q), out, cfl, bcl, bcu, [1, 2], meth, [4, 4], dx, dy,
                  dt, aux=_dev(aux), aux_bc_lower=bcl, aux_bc_upper=bcu)
    torch.cuda.synchronize()
    got = q_from_device(out)
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    assert float(cfl.item()) == max(c1, c2)


def test_api_acoustics_2d_variable_io_fixture(golden_dir, tmp_path):
    """examples/acoustics_2d_variable through the PyClaw API with ascii output:
    test_acoustics_2d_variable_io.py (frames 0 and 20 written, read back and
    compared with io_test_verification at relative 1e-6, aux included) and
    test_acoustics_2d_variable.py (pressure, abstol 0.1); the run equals the
    oracle's bit for bit."""
    import pyclaw_amd as pyclaw
    from pyclaw_amd import riemann
    from tests.test_oracle import vc_acoustics_run
    q0, aux0, d = C.acoustics_2d_variable()
    solver = pyclaw.ClawSolver2D(riemann.vc_acoustics_2D)
    solver.dimensional_split = False
    solver.limiters = pyclaw.limiters.tvd.MC
    solver.bc_lower[0] = pyclaw.BC.wall
    solver.bc_upper[0] = pyclaw.BC.extrap
    solver.bc_lower[1] = pyclaw.BC.wall
    solver.bc_upper[1] = pyclaw.BC.extrap
    solver.aux_bc_lower[0] = pyclaw.BC.wall
    solver.aux_bc_upper[0] = pyclaw.BC.extrap
    solver.aux_bc_lower[1] = pyclaw.BC.wall
    solver.aux_bc_upper[1] = pyclaw.BC.extrap
    x = pyclaw.Dimension(-1.0, 1.0, 50, name="x")
    y = pyclaw.Dimension(-1.0, 1.0, 50, name="y")
    domain = pyclaw.Domain([x, y])
    state = pyclaw.State(domain, 3, 2)
    state.aux[...] = aux0
    state.q[...] = q0
    claw = pyclaw.Controller()
    claw.keep_copy = True
    claw.solution = pyclaw.Solution(state, domain)
    claw.solver = solver
    claw.outdir = str(tmp_path)
    claw.tfinal = 0.6
    claw.num_output_times = 20
    claw.write_aux_init = True
    claw.run()
    ver = f"{golden_dir}/acoustics_2d_variable_io"
    exp_aux = pyclaw.Solution(0, path=ver, read_aux=True).state.aux
    exp_q = pyclaw.Solution(20, path=ver, read_aux=False).state.q
    got_aux = pyclaw.Solution(0, path=str(tmp_path), read_aux=True).state.get_aux_global()
    got_q = pyclaw.Solution(20, path=str(tmp_path), read_aux=False).state.get_q_global()
    rel_q = np.linalg.norm(exp_q - got_q) / np.linalg.norm(exp_q)
    rel_aux = np.linalg.norm(exp_aux - got_aux) / np.linalg.norm(exp_aux)
    print("variable acoustics io: q", rel_q, "aux", rel_aux)
    assert rel_q < 1e-6 and rel_aux < 1e-6
    p = np.load(f"{golden_dir}/acoustics_2d_variable_pressure_classic.npz")["pressure"]
    final = claw.frames[-1].state.q
    assert np.abs(p.reshape(-1) - final[0].reshape(-1)).max() < 1e-1
    run = vc_acoustics_run(q0, aux0, d)
    run.run(0.6, 20)
    assert claw.solver.status["numsteps"] == run.numsteps
    assert np.array_equal(final, run.q), np.abs(final - run.q).max()
