"""Aux fields on the device: the dimension-split 2D step with the variable-
coefficient advection solver of examples/advection_reaction_2d (edge velocities
in two aux fields), against the oracle (whose aux step equals the reference's
own Fortran step2ds with that example's rpn2, tests/test_oracle.py), through the
PyClaw API with before_step / step_source callbacks (the example's fixture), and
through the fmod drop-in."""
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
    """A solver with aux fields needs them; the unsplit step does not take it."""
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
