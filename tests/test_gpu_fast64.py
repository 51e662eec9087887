"""The FMA-contracted FP64 variant (dtype CLAWHIP_F64_FAST, precision
"fp64fast"): a step of every kernel family it reaches -- the classic unsplit and
dimension-split 2D steps, the SharpClaw operator with component and
characteristic WENO5 -- stays within 1e-12 of the oracle's step (relative to the
state's, or the operator's, largest magnitude), with the CFL number within 1e-12
relative; and in at least one case it is not bit-equal (the contracted kernels
are the ones that ran).  Whole runs of the same solver in the variant match
between the device-controlled and host-driven loops bit for bit elsewhere
(test_gpu_sharp_device.py, test_gpu_ensemble.py, test_gpu_configs.py)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests._cases import seeded_state

pytestmark = pytest.mark.gpu

BAR = 1e-12

PD = {O.ACOUSTICS: dict(rho=1.0, bulk=4.0), O.EULER: dict(gamma=1.4),
      O.SHALLOW: dict(grav=1.0), O.ADVECTION: dict(u=0.5, v=-0.7)}


def _orp(kind):
    pd = PD[kind]
    return O.make_rp(kind, efix=1, rho=pd.get("rho", 1.0), bulk=pd.get("bulk", 4.0),
                     gamma=pd.get("gamma", 1.4), grav=pd.get("grav", 1.0),
                     u=pd.get("u", 0.5), v=pd.get("v", 1.0))


def _close(got, ref, scale=None):
    scale = np.abs(ref).max() if scale is None else scale
    return np.abs(got - ref).max() / scale


def _mild_euler(q):
    """Cell-to-cell noise of the full admissible range can make WENO5 edge
    pressures negative; the operator tests keep it milder."""
    return 0.85 * seeded_state(np.random.default_rng(0), O.EULER, (1, 1)).mean(
        axis=(1, 2))[:, None, None] + 0.15 * q


@pytest.mark.parametrize("kind", [O.ACOUSTICS, O.EULER, O.SHALLOW, O.ADVECTION])
def test_unsplit_step_fast64_within_bar(kind):
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device, q_to_device
    print("fast64 unsplit", kind, flush=True)
    rng = np.random.default_rng(300 + kind)
    mw = O.MWAVES[kind]
    differs = False
    for (mx, my), (bcl, bcu), (trans, lim) in zip(
            [(67, 35), (130, 17), (33, 64)], [([1, 1], [1, 1]), ([2, 2], [2, 2]), ([3, 1], [2, 3])],
            [(2, 4), (1, 3), (2, 1)]):
        if kind == O.ADVECTION and (3 in bcl or 3 in bcu):
            bcl, bcu = [2, 1], [2, 1]
        q = seeded_state(rng, kind, (mx, my))
        method = O.method_array(order=2, transverse=trans)
        dx, dy = 1.0 / mx, 0.8 / my
        dt = 0.3 * min(dx, dy)
        ng = 2
        qbc = O.with_ghosts(q, ng)
        O.bc2(qbc, ng, bcl, bcu)
        qnew = qbc.copy(order="F")
        cref = O.step2(_orp(kind), qbc, qnew, ng, dx, dy, dt, method, [lim] * mw, ids=0)
        ref = O.interior(qnew, ng)
        qi = q_to_device(q)
        qo = torch.empty_like(qi)
        cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
        K.step2_unsplit(K.make_rp(kind, PD[kind]), qi, qo, cfl, bcl, bcu, [1, 2], method,
                        [lim] * mw, dx, dy, dt, fast=True)
        torch.cuda.synchronize()
        got = q_from_device(qo)
        err = _close(got, ref)
        assert err <= BAR, ((mx, my), err)
        assert abs(float(cfl.item()) - cref) <= BAR * cref
        differs = differs or not np.array_equal(got, ref)
    if kind in (O.EULER, O.SHALLOW):
        assert differs   # the contracted FP64 kernels ran


@pytest.mark.parametrize("kind", [O.EULER, O.SHALLOW, O.ACOUSTICS])
def test_split_step_fast64_within_bar(kind):
    """step2ds x then y (classic/step2ds.f90) as ClawSolver2D takes them."""
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device, q_to_device
    print("fast64 split", kind, flush=True)
    rng = np.random.default_rng(400 + kind)
    mw = O.MWAVES[kind]
    mx, my, ng = 77, 45, 2
    q = seeded_state(rng, kind, (mx, my))
    method = O.method_array(dimensional_split=True)
    bcl, bcu = [3, 1], [2, 3]
    dx, dy = 1.0 / mx, 0.8 / my
    dt = 0.3 * min(dx, dy)
    qbc = O.with_ghosts(q, ng)
    O.bc2(qbc, ng, bcl, bcu)
    qold = qbc.copy(order="F")
    c1 = O.step2(_orp(kind), qold, qbc, ng, dx, dy, dt, method, [4] * mw, ids=1)
    c2 = O.step2(_orp(kind), qbc, qbc, ng, dx, dy, dt, method, [4] * mw, ids=2)
    ref = O.interior(qbc, ng)
    qi = q_to_device(q)
    qo = torch.empty_like(qi)
    cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
    K.step2_split(K.make_rp(kind, PD[kind]), qi, qo, cfl, bcl, bcu, [1, 2], method, [4] * mw,
                  dx, dy, dt, fast=True)
    torch.cuda.synchronize()
    err = _close(q_from_device(qo), ref)
    assert err <= BAR, err
    assert abs(float(cfl.item()) - max(c1, c2)) <= BAR * max(c1, c2)


def test_split3_step_fast64_within_bar():
    """step3ds x, y, z (classic/step3ds.f90) with euler_3D in the contracted
    variant (its shared reciprocals included): within 1e-12 of the oracle."""
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device, q_to_device
    print("fast64 split 3D", flush=True)
    rng = np.random.default_rng(600)
    mx, my, mz, ng = 23, 17, 13, 2
    q = seeded_state(rng, O.EULER3, (mx, my, mz))
    meth = O.method_array(dimensional_split=True)
    bcl, bcu = [1, 3, 2], [3, 1, 2]
    dx, dy, dz = 1.0 / mx, 0.9 / my, 1.1 / mz
    dt = 0.2 * min(dx, dy, dz)
    qbc = O.with_ghosts(q, ng)
    O.bc3(qbc, ng, bcl, bcu)
    qold = qbc.copy(order="F")
    rp = O.make_rp(O.EULER3)
    cs = [O.step3ds(rp, qold, qbc, ng, dx, dy, dz, dt, meth, [4] * 3, 1),
          O.step3ds(rp, qbc, qbc, ng, dx, dy, dz, dt, meth, [4] * 3, 2),
          O.step3ds(rp, qbc, qbc, ng, dx, dy, dz, dt, meth, [4] * 3, 3)]
    ref = O.interior(qbc, ng)
    qi = q_to_device(q)
    qo = torch.empty_like(qi)
    cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
    K.step3_split(K.make_rp(5, {"gamma": 1.4}), qi, qo, cfl, bcl, bcu, [1, 2, 3], meth,
                  [4] * 3, dx, dy, dz, dt, fast=True)
    torch.cuda.synchronize()
    got = q_from_device(qo)
    assert _close(got, ref) <= BAR, _close(got, ref)
    assert abs(float(cfl.item()) - max(cs)) <= BAR * max(cs)
    assert not np.array_equal(got, ref)


@pytest.mark.parametrize("kind,cd", [(O.SHALLOW, 2), (O.EULER, 2), (O.EULER, 0)])
def test_sharp_operator_fast64_within_bar(kind, cd):
    """dq = dt*L(q) (sharpclaw/flux2.f90) with WENO5 in the contracted variant:
    within 1e-12 of the oracle's, relative to the largest |dq|."""
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device, q_to_device
    print("fast64 sharp", kind, cd, flush=True)
    rng = np.random.default_rng(500 + 10 * kind + cd)
    mx, my, ng = 45, 37, 3
    q = seeded_state(rng, kind, (mx, my))
    if kind == O.EULER:
        q = _mild_euler(q)
    bcl, bcu = [1, 3], [3, 1]
    qbc = O.with_ghosts(q, ng)
    O.bc2(qbc, ng, bcl, bcu)
    dx, dy = 1.0 / mx, 0.9 / my
    dt = 0.2 * min(dx, dy)
    dq, cref = O.sharp_flux2(_orp(kind), qbc, ng, dx, dy, dt, cd)
    ref = O.interior(dq, ng)
    qi = q_to_device(q)
    out = torch.empty_like(qi)
    cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
    K.sharp2_stage(K.make_rp(kind, PD[kind]), cd, out, cfl, bcl, bcu, [1, 2], dt, dx, dy, qi,
                   fast=True)
    torch.cuda.synchronize()
    got = q_from_device(out)
    err = _close(got, ref)
    assert err <= BAR, err
    assert abs(float(cfl.item()) - cref) <= BAR * cref
    assert not np.array_equal(got, ref)
