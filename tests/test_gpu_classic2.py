"""GPU parity of the 2D classic step against the oracle (FP64: bit for bit).

Inputs are seeded admissible states on small grids (sizes chosen so strips and
row blocks are ragged); the oracle fills ghost cells with the reference's BC
rules and takes the reference's step (oracle == reference Fortran bitwise, see
test_oracle.py); the HIP kernel takes the same step from the compact state with
its own boundary mapping.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests._cases import seeded_state

pytestmark = pytest.mark.gpu

K = None


def _k():
    global K
    if K is None:
        from pyclaw_amd import kernels
        K = kernels
    return K


KIND_PD = {
    O.ACOUSTICS: dict(rho=1.0, bulk=4.0),
    O.EULER: dict(gamma=1.4),
    O.SHALLOW: dict(grav=1.0),
    O.ADVECTION: dict(u=0.5, v=-0.7),
}


def _oracle_rp(kind):
    pd = KIND_PD[kind]
    return O.make_rp(kind, efix=1, rho=pd.get("rho", 1.0), bulk=pd.get("bulk", 4.0),
                     gamma=pd.get("gamma", 1.4), grav=pd.get("grav", 1.0),
                     u=pd.get("u", 0.5), v=pd.get("v", 1.0))


def _gpu_step(kind, q, bcl, bcu, method, mthlim, dx, dy, dt, dtype=torch.float64):
    from pyclaw_amd.device import q_to_device, q_from_device
    k = _k()
    rp = k.make_rp(kind, KIND_PD[kind])
    qi = q_to_device(q).to(dtype)
    qo = torch.empty_like(qi)
    cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
    comp = torch.zeros_like(qi) if dtype == torch.float32 else None
    k.step2_unsplit(rp, qi, qo, cfl, bcl, bcu, [1, 2], method, mthlim, dx, dy, dt, comp_in=comp,
                    comp_out=comp)
    torch.cuda.synchronize()
    return q_from_device(qo), float(cfl.item())


def _oracle_step(kind, q, bcl, bcu, method, mthlim, dx, dy, dt):
    ng = 2
    qbc = O.with_ghosts(q, ng)
    O.bc2(qbc, ng, bcl, bcu)
    qnew = qbc.copy(order="F")
    cfl = O.step2(_oracle_rp(kind), qbc, qnew, ng, dx, dy, dt, method, mthlim, ids=0)
    return O.interior(qnew, ng), cfl


BCS = [([1, 1], [1, 1]), ([2, 2], [2, 2]), ([3, 1], [2, 3]), ([3, 3], [3, 3])]


@pytest.mark.parametrize("kind", [O.ACOUSTICS, O.EULER, O.SHALLOW, O.ADVECTION])
@pytest.mark.parametrize("trans,lim,order", [(2, 4, 2), (1, 3, 2), (0, 1, 2), (2, 2, 1), (2, 5, 2)])
def test_unsplit_step_bit_exact(kind, trans, lim, order):
    rng = np.random.default_rng(100 * kind + 10 * trans + lim)
    mw = O.MWAVES[kind]
    for (mx, my), (bcl, bcu) in zip([(67, 35), (130, 17), (5, 3), (121, 64)], BCS):
        if kind == O.ADVECTION and (3 in bcl or 3 in bcu):
            bcl, bcu = [2, 1], [2, 1]
        q = seeded_state(rng, kind, (mx, my))
        method = O.method_array(order=order, transverse=trans)
        dx, dy = 1.0 / mx, 0.8 / my
        dt = 0.3 * min(dx, dy)
        ref, cref = _oracle_step(kind, q, bcl, bcu, method, [lim] * mw, dx, dy, dt)
        out, cgpu = _gpu_step(kind, q, bcl, bcu, method, [lim] * mw, dx, dy, dt)
        assert np.array_equal(out, ref), (mx, my, bcl, bcu, np.abs(out - ref).max())
        assert cgpu == cref, (cgpu, cref)


@pytest.mark.parametrize("kind", [O.EULER, O.SHALLOW])
def test_unsplit_fp32_kahan_close(kind):
    rng = np.random.default_rng(5)
    q = seeded_state(rng, kind, (96, 70))
    method = O.method_array(transverse=2)
    mw = O.MWAVES[kind]
    ref, cref = _oracle_step(kind, q, [1, 2], [1, 2], method, [4] * mw, 0.01, 0.01, 0.003)
    out, cgpu = _gpu_step(kind, q, [1, 2], [1, 2], method, [4] * mw, 0.01, 0.01, 0.003,
                          dtype=torch.float32)
    assert np.abs(out - ref).max() / np.abs(ref).max() < 2e-6
    assert abs(cgpu - cref) / cref < 1e-6
