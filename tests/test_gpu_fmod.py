"""The `solver.fmod` drop-in modules called exactly as the reference's solvers
call their f2py kernels (classic/solver.py:527-541, 680-690; sharpclaw/solver.py:
893), on ghost-filled F-order qbc arrays: bitwise equal to the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from tests._cases import seeded_state

pytestmark = pytest.mark.gpu

PD = {O.ACOUSTICS: dict(rho=1.0, bulk=4.0), O.EULER: dict(gamma=1.4),
      O.SHALLOW: dict(grav=1.0), O.ADVECTION: dict(u=0.5, v=-0.7), O.EULER3: dict(gamma=1.4)}


def _orp(kind):
    pd = PD[kind]
    return O.make_rp(kind, efix=1, rho=pd.get("rho", 1.0), bulk=pd.get("bulk", 4.0),
                     gamma=pd.get("gamma", 1.4), grav=pd.get("grav", 1.0),
                     u=pd.get("u", 0.5), v=pd.get("v", 1.0))


def _fmod(name, kind):
    import importlib
    mod = importlib.reload(importlib.import_module("pyclaw_amd.fmod." + name))
    for k, v in PD[kind].items():          # State.set_cparam (state.py:220-231)
        setattr(mod.cparam, k, v)
    mod.cparam.efix = 1
    return mod


@pytest.mark.parametrize("kind", [O.ACOUSTICS, O.EULER, O.SHALLOW, O.ADVECTION])
def test_fmod_step2_and_step2ds(kind):
    fm = _fmod("classic2", kind)
    rng = np.random.default_rng(60 + kind)
    mx, my, ng = 45, 31, 2
    q = seeded_state(rng, kind, (mx, my))
    bcl, bcu = [1, 3], [2, 2]
    if kind == O.ADVECTION:
        bcl, bcu = [2, 1], [2, 1]
    qbc = O.with_ghosts(q, ng)
    O.bc2(qbc, ng, bcl, bcu)
    dx, dy, dt = 1.0 / mx, 0.8 / my, 0.002
    mw = O.MWAVES[kind]
    mthlim = np.array([4] * mw)
    # unsplit
    meth = np.array(O.method_array(transverse=2))
    ref = qbc.copy(order="F")
    cref = O.step2(_orp(kind), qbc, ref, ng, dx, dy, dt, list(meth), list(mthlim))
    qnew = qbc.copy(order="F")
    qnew, cfl = fm.step2(max(mx, my), ng, mx, my, qbc.copy(order="F"), qnew, None, dx, dy, dt,
                         meth, mthlim, None, None, None, None, False, kind, kind)
    assert np.array_equal(O.interior(qnew, ng), O.interior(ref, ng))
    assert cfl == cref
    # dimension split: sweep 1 (every cell) then sweep 2 in place
    meth = np.array(O.method_array(dimensional_split=True))
    r1 = qbc.copy(order="F")
    c1 = O.step2(_orp(kind), qbc, r1, ng, dx, dy, dt, list(meth), list(mthlim), ids=1)
    r2 = r1.copy(order="F")
    c2 = O.step2(_orp(kind), r1, r2, ng, dx, dy, dt, list(meth), list(mthlim), ids=2)
    qb = qbc.copy(order="F")
    qb, g1 = fm.step2ds(max(mx, my), ng, mx, my, qbc.copy(order="F"), qb, None, dx, dy, dt,
                        meth, mthlim, None, None, None, None, 1, False, kind, kind)
    assert np.array_equal(qb, r1)
    assert g1 == c1
    qb, g2 = fm.step2ds(max(mx, my), ng, mx, my, qb, qb, None, dx, dy, dt, meth, mthlim, None,
                        None, None, None, 2, False, kind, kind)
    assert np.array_equal(O.interior(qb, ng), O.interior(r2, ng))
    assert g2 == c2


def test_fmod_step3ds():
    kind = O.EULER3
    fm = _fmod("classic3", kind)
    rng = np.random.default_rng(70)
    mx, my, mz, ng = 21, 14, 9, 2
    q = seeded_state(rng, kind, (mx, my, mz))
    qbc = O.with_ghosts(q, ng)
    O.bc3(qbc, ng, [1, 3, 2], [3, 1, 2])
    d = (1.0 / mx, 1.0 / my, 1.0 / mz)
    dt = 0.002
    meth = np.array(O.method_array(dimensional_split=True))
    mthlim = np.array([4, 4, 4])
    ref = qbc.copy(order="F")
    got = qbc.copy(order="F")
    src = qbc.copy(order="F")
    for idir in (1, 2, 3):
        nxt = ref.copy(order="F")
        cref = O.step3ds(_orp(kind), ref, nxt, ng, *d, dt, list(meth), list(mthlim), idir)
        ref = nxt
        got, cfl = fm.step3ds(max(mx, my, mz), ng, mx, my, mz, src, got, None, *d, dt, meth,
                              mthlim, None, None, None, None, idir, False, kind, kind, kind)
        src = got
        assert cfl == cref, idir
    assert np.array_equal(O.interior(got, ng), O.interior(ref, ng))


def test_fmod_step3():
    """fmod.step3 (the unsplit 3D step, trans_cor) on a filled qbc: the interior
    equals the oracle's step3 bit for bit, as does the CFL."""
    kind = O.EULER3
    fm = _fmod("classic3", kind)
    rng = np.random.default_rng(71)
    mx, my, mz, ng = 17, 12, 10, 2
    q = seeded_state(rng, kind, (mx, my, mz))
    qbc = O.with_ghosts(q, ng)
    O.bc3(qbc, ng, [3, 1, 2], [1, 3, 2])
    d = (1.0 / mx, 1.0 / my, 1.0 / mz)
    meth = np.array([1, 2, 22, 0, 0, 0, 0])
    mthlim = np.array([1, 1, 1])
    ref = qbc.copy(order="F")
    cref = O.step3(_orp(kind), qbc, ref, ng, *d, 0.003, list(meth), list(mthlim))
    got, cfl = fm.step3(max(mx, my, mz), ng, mx, my, mz, qbc, qbc.copy(order="F"), None, *d,
                        0.003, meth, mthlim, None, None, None, None, False, kind, kind, kind)
    assert np.array_equal(O.interior(got, ng), O.interior(ref, ng)) and cfl == cref


@pytest.mark.parametrize("kind,cd", [(O.SHALLOW, 2), (O.EULER, 2), (O.ACOUSTICS, 0)])
def test_fmod_sharpclaw_flux2(kind, cd):
    fm = _fmod("sharpclaw2", kind)
    p = fm.clawparams
    p.char_decomp = cd
    p.num_waves = O.MWAVES[kind]
    p.alloc_clawparams()
    rng = np.random.default_rng(80 + kind)
    mx, my, ng = 40, 27, 3
    q = seeded_state(rng, kind, (mx, my))
    qbc = O.with_ghosts(q, ng)
    O.bc2(qbc, ng, [3, 2], [1, 2])
    p.dx = [1.0 / mx, 1.0 / my]
    dt = 0.003
    dq_ref, cref = O.sharp_flux2(_orp(kind), qbc, ng, p.dx[0], p.dx[1], dt, cd)
    dq, cfl = fm.flux2(qbc, None, dt, 0.0, ng, max(mx, my), mx, my, kind, None)
    assert np.array_equal(O.interior(dq, ng), O.interior(dq_ref, ng))
    assert cfl == cref
