"""The unsplit 3D step on the device (clawhip_step3_unsplit, classic/step3.f90
with flux3.f90's transverse and double-transverse terms): bitwise against the
oracle, whose step3 equals the reference's own Fortran step3 on every qbc cell
(tests/test_oracle.py), for every transverse variant; with stored ghosts; and
examples/euler_3d/shock_bubble.py through the PyClaw API (ClawSolver3D's
unsplit trans_cor default, the incoming-shock custom BC) bitwise against the
oracle's run and within the fixture's bar."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests import _cases as C
from tests._cases import seeded_state

pytestmark = pytest.mark.gpu

CASES = [((19, 13, 11), [1, 3, 2], [3, 1, 2], 22, 2, 4),
         ((21, 7, 9), [2, 2, 3], [2, 2, 1], 21, 2, 1),
         ((9, 17, 8), [3, 1, 1], [1, 3, 3], 20, 2, 4),
         ((12, 10, 14), [1, 1, 1], [1, 1, 1], 11, 2, 1),
         ((8, 9, 10), [3, 3, 3], [3, 3, 3], 11, 1, 4),
         ((10, 8, 9), [2, 1, 3], [2, 3, 1], 10, 2, 4),
         ((7, 6, 5), [1, 2, 1], [1, 2, 1], 0, 2, 4)]


@pytest.mark.parametrize("shape,bcl,bcu,trans,order,lim", CASES)
def test_step3_bit_exact(shape, bcl, bcu, trans, order, lim):
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device, q_to_device
    print("step3", shape, trans, order, lim, flush=True)
    rng = np.random.default_rng(sum(shape) + trans)
    q = seeded_state(rng, O.EULER3, shape)
    ng = 2
    dx, dy, dz = 1.0 / shape[0], 0.9 / shape[1], 1.1 / shape[2]
    dt = 0.15 * min(dx, dy, dz)
    meth = [1, order, trans, 0, 0, 0, 0]
    qbc = O.with_ghosts(q, ng)
    O.bc3(qbc, ng, bcl, bcu)
    ref = qbc.copy(order="F")
    c_ref = O.step3(O.make_rp(O.EULER3), qbc, ref, ng, dx, dy, dz, dt, meth, [lim] * 3)
    rp = K.make_rp(5, {"gamma": 1.4})
    qi = q_to_device(q)
    out = torch.empty_like(qi)
    cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
    K.step3_unsplit(rp, qi, out, cfl, bcl, bcu, [1, 2, 3], meth, [lim] * 3, dx, dy, dz, dt)
    torch.cuda.synchronize()
    got = q_from_device(out)
    want = O.interior(ref, ng)
    assert np.array_equal(got, want), np.abs(got - want).max()
    assert float(cfl.item()) == c_ref


def test_step3_stored_ghosts_bit_exact():
    """A caller-filled qbc (BC code 4) gives the same step as the BC fill."""
    from pyclaw_amd import _lib as L
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device, to_planar
    print("step3 stored ghosts", flush=True)
    rng = np.random.default_rng(77)
    shape, ng = (13, 11, 9), 2
    q = seeded_state(rng, O.EULER3, shape)
    qbc = O.with_ghosts(q, ng)
    O.bc3(qbc, ng, [3, 1, 3], [1, 3, 1])
    qbc[:, :ng] = 1.1 * qbc[:, ng:2 * ng]      # a user's ghost state (physical: scaled)
    dx, dy, dz = 1.0 / 13, 1.0 / 11, 1.0 / 9
    dt = 0.01
    meth = [1, 2, 22, 0, 0, 0, 0]
    ref = qbc.copy(order="F")
    c_ref = O.step3(O.make_rp(O.EULER3), qbc, ref, ng, dx, dy, dz, dt, meth, [4] * 3)
    qi = torch.from_numpy(to_planar(qbc)).cuda()
    out = torch.empty((5,) + shape[::-1], dtype=torch.float64, device="cuda")
    cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
    K.step3_unsplit(K.make_rp(5, {"gamma": 1.4}), qi, out, cfl, [4] * 3, [4] * 3, [1, 2, 3],
                    meth, [4] * 3, dx, dy, dz, dt, gin=L.geom3(*shape, ng))
    torch.cuda.synchronize()
    assert np.array_equal(q_from_device(out), O.interior(ref, ng))
    assert float(cfl.item()) == c_ref


def test_api_shock_bubble_3d_fixture(golden_dir):
    """examples/euler_3d/test_euler3d.py:31-34 through the API: ClawSolver3D
    with dimensional_split=False (trans_cor, minmod), the incoming-shock BC
    function, 64 x 16 x 16 to t = 0.1 with 10 outputs: every output frame
    bitwise equal to the oracle's run, density L1 below 1e-6 of the fixture."""
    import pyclaw_amd as pyclaw
    from pyclaw_amd import riemann
    from tests.test_oracle import shock_bubble_3d_run
    print("shock bubble 3D through the API", flush=True)
    q, d, inc = C.shock_bubble_3d()

    def incoming_shock(state, dim, t, qbc, auxbc, num_ghost):
        for i in range(num_ghost):
            qbc[:, i, ...] = inc[:, None, None]

    solver = pyclaw.ClawSolver3D(riemann.euler_3D)
    solver.dimensional_split = False
    solver.all_bcs = pyclaw.BC.extrap
    solver.bc_lower[0] = pyclaw.BC.custom
    solver.user_bc_lower = incoming_shock
    solver.bc_lower[1] = pyclaw.BC.wall
    solver.bc_lower[2] = pyclaw.BC.wall
    x = pyclaw.Dimension(0.0, 2.0, 64, name="x")
    y = pyclaw.Dimension(0.0, 0.5, 16, name="y")
    z = pyclaw.Dimension(0.0, 0.5, 16, name="z")
    domain = pyclaw.Domain([x, y, z])
    state = pyclaw.State(domain, 5)
    state.problem_data["gamma"] = 1.4
    state.q[...] = q
    claw = pyclaw.Controller()
    claw.solution = pyclaw.Solution(state, domain)
    claw.solver = solver
    claw.output_format = None
    claw.keep_copy = True
    claw.tfinal = 0.1
    claw.num_output_times = 10
    claw.run()
    run, frames, _ = shock_bubble_3d_run()
    assert claw.solver.status["numsteps"] == run.numsteps
    for k in range(11):
        assert np.array_equal(claw.frames[k].state.q, frames[k]), k
    exp = np.load(f"{golden_dir}/euler_3d_expected.npz")["shock_bubble"]
    err = d[0] * d[1] * d[2] * np.abs(claw.frames[10].state.q[0].reshape(-1) - exp).sum()
    print("shock bubble 3D L1", err)
    assert err < 1e-6
