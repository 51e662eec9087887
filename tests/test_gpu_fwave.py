"""f-waves on the device (classic/flux2.f90:134-160 with use_fwave): the
second-order correction with dsign(1, s) in place of |s|, in the split and
unsplit steps and the fmod sweep, and the shallow_bathymetry_fwave_2D solver --
each bitwise against the oracle, whose f-wave steps equal the reference's own
Fortran step2 / step2ds (tests/test_oracle.py); and the examples/shallow_2d
sill run through the PyClaw API, bitwise against the oracle's run and within
its fixture's bar."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests import _cases as C
from tests._cases import seeded_state

pytestmark = pytest.mark.gpu


def _dev(q, dtype=0):
    from pyclaw_amd.device import q_to_device
    return q_to_device(q, dtype)


def _swe_bathymetry(rng, mx, my):
    h = rng.uniform(0.5, 1.5, (mx, my))
    q = np.asfortranarray(np.array([h, h * rng.uniform(-0.6, 0.6, (mx, my)),
                                     h * rng.uniform(-0.6, 0.6, (mx, my))]))
    aux = np.asfortranarray(rng.uniform(0.0, 0.4, (1, mx, my)))
    return q, aux


@pytest.mark.parametrize("bcl,bcu", [([3, 1], [1, 3]), ([2, 3], [2, 1])])
def test_split_step_fwave_bathymetry_bit_exact(bcl, bcu):
    """Both sweeps with shallow_bathymetry_fwave_2D, compact output; FP32 +
    Kahan within 1e-5 of FP64."""
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device
    print("f-wave split", bcl, bcu, flush=True)
    rng = np.random.default_rng(41)
    mx, my, ng = 67, 43, 2
    q, aux = _swe_bathymetry(rng, mx, my)
    dx, dy, dt = 1.0 / mx, 1.0 / my, 0.002
    meth = O.method_array(dimensional_split=True, num_aux=1)
    qbc, abc = O.with_ghosts(q, ng), O.with_ghosts(aux, ng)
    O.bc2(qbc, ng, bcl, bcu)
    O.bc2(abc, ng, [1, 1], [1, 1], reflect=(-1, -1))
    rp_o = O.make_rp(O.SHALLOW_FWAVE, grav=1.0)
    qold = qbc.copy(order="F")
    c1 = O.step2(rp_o, qold, qbc, ng, dx, dy, dt, meth, [1] * 3, ids=1, use_fwave=True, aux=abc)
    c2 = O.step2(rp_o, qbc, qbc, ng, dx, dy, dt, meth, [1] * 3, ids=2, use_fwave=True, aux=abc)
    ref = O.interior(qbc, ng)
    pd = {"grav": 1.0, "dry_tolerance": 1e-3}
    rp = K.make_rp(11, pd, fwave=True)
    out = torch.empty_like(_dev(q))
    cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
    K.step2_split(rp, _dev(q), out, cfl, bcl, bcu, [1, 2], meth, [1] * 3, dx, dy, dt,
                  aux=_dev(aux), aux_bc_lower=[1, 1], aux_bc_upper=[1, 1])
    torch.cuda.synchronize()
    got = q_from_device(out)
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    assert float(cfl.item()) == max(c1, c2)
    q32 = _dev(q, 1)
    out32, comp_out = torch.empty_like(q32), torch.empty_like(q32)
    cfl32 = torch.zeros(1, dtype=torch.float64, device="cuda")
    K.step2_split(rp, q32, out32, cfl32, bcl, bcu, [1, 2], meth, [1] * 3, dx, dy, dt,
                  comp_in=torch.zeros_like(q32), comp_out=comp_out, aux=_dev(aux, 1),
                  aux_bc_lower=[1, 1], aux_bc_upper=[1, 1])
    torch.cuda.synchronize()
    assert np.abs(q_from_device(out32) - ref).max() <= 1e-5 * np.abs(ref).max()


def _lake_with_island(rng, mx, my):
    """A lake at rest (surface 1) over random bathymetry, a dry island (zero
    depth on ground at 1.2) and a shore ring of cells below dry_tolerance 1e-3
    that hold momentum the solver ignores; moving water away from the island."""
    X, Y = np.meshgrid(np.arange(mx), np.arange(my), indexing="ij")
    b = rng.uniform(0.0, 0.4, (mx, my))
    r = np.hypot(X - 16, Y - 11)
    island, shore = r < 5.0, (r >= 5.0) & (r < 6.5)
    b[island] = 1.2
    b[shore] = rng.uniform(0.9991, 0.99995, shore.sum())
    h = np.maximum(1.0 - b, 0.0)
    h[island] = 0.0
    u, v = np.zeros((mx, my)), np.zeros((mx, my))
    far = X > 32
    u[far], v[far] = rng.uniform(-0.4, 0.4, far.sum()), rng.uniform(-0.4, 0.4, far.sum())
    hu, hv = h * u, h * v
    hu[shore] = rng.uniform(-2e-4, 2e-4, shore.sum())
    hv[shore] = rng.uniform(-2e-4, 2e-4, shore.sum())
    return np.asfortranarray(np.array([h, hu, hv])), np.asfortranarray(b[None])


def test_split_step_fwave_dry_cells_bit_exact():
    """shallow_bathymetry_fwave_2D where cells are dry (_lake_with_island):
    interfaces between two dry cells, between a dry cell holding momentum and a
    wet one, and between wet cells.  Both sweeps in FP64 equal the oracle's,
    whose dry branch is pinned by properties in tests/test_oracle.py (no waves
    between two dry cells, a dry cell's momentum not entering, the
    flux-difference sum, x -> -x symmetry); the mistakes planted there (13, 14,
    15) each change this step's result."""
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device
    print("f-wave split, dry cells", flush=True)
    mx, my, ng = 53, 41, 2
    q, aux = _lake_with_island(np.random.default_rng(47), mx, my)
    dx, dy, dt = 1.0 / mx, 1.0 / my, 0.002
    meth = O.method_array(dimensional_split=True, num_aux=1)
    bcl, bcu = [1, 3], [3, 1]
    qbc, abc = O.with_ghosts(q, ng), O.with_ghosts(aux, ng)
    O.bc2(qbc, ng, bcl, bcu)
    O.bc2(abc, ng, [1, 1], [1, 1], reflect=(-1, -1))
    rp_o = O.make_rp(O.SHALLOW_FWAVE, grav=1.0, dry_tolerance=1e-3)
    qold = qbc.copy(order="F")
    c1 = O.step2(rp_o, qold, qbc, ng, dx, dy, dt, meth, [4] * 3, ids=1, use_fwave=True, aux=abc)
    c2 = O.step2(rp_o, qbc, qbc, ng, dx, dy, dt, meth, [4] * 3, ids=2, use_fwave=True, aux=abc)
    ref = O.interior(qbc, ng)
    assert np.isfinite(ref).all() and (ref[0] <= 1e-3).sum() > 100
    rp = K.make_rp(11, {"grav": 1.0, "dry_tolerance": 1e-3}, fwave=True)
    out = torch.empty_like(_dev(q))
    cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
    K.step2_split(rp, _dev(q), out, cfl, bcl, bcu, [1, 2], meth, [4] * 3, dx, dy, dt,
                  aux=_dev(aux), aux_bc_lower=[1, 1], aux_bc_upper=[1, 1])
    torch.cuda.synchronize()
    got = q_from_device(out)
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    assert float(cfl.item()) == max(c1, c2)


def test_split_step_whole_qbc_output_bit_exact():
    """A stored-ghost input and the whole-qbc output: every cell, ghosts
    included, equals the oracle's qbc after its two in-place sweeps (the x-sweep
    writes the ghost rows, the y-sweep the ghost columns, corners stay)."""
    from pyclaw_amd import _lib as L
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import from_planar, to_planar
    print("f-wave split, whole qbc", flush=True)
    rng = np.random.default_rng(43)
    mx, my, ng = 37, 29, 2
    q, aux = _swe_bathymetry(rng, mx, my)
    qbc, abc = O.with_ghosts(q, ng), O.with_ghosts(aux, ng)
    O.bc2(qbc, ng, [3, 1], [1, 3])
    qbc[:, :ng, :] = rng.uniform(0.5, 1.0, (3, ng, my + 2 * ng))   # a user's ghost values
    O.bc2(abc, ng, [1, 1], [1, 1], reflect=(-1, -1))
    dx, dy, dt = 1.0 / mx, 1.0 / my, 0.002
    meth = O.method_array(dimensional_split=True, num_aux=1)
    rp_o = O.make_rp(O.SHALLOW_FWAVE, grav=1.0)
    ref = qbc.copy(order="F")
    qold = ref.copy(order="F")
    O.step2(rp_o, qold, ref, ng, dx, dy, dt, meth, [4] * 3, ids=1, use_fwave=True, aux=abc)
    O.step2(rp_o, ref, ref, ng, dx, dy, dt, meth, [4] * 3, ids=2, use_fwave=True, aux=abc)
    rp = K.make_rp(11, {"grav": 1.0}, fwave=True)
    qin = torch.from_numpy(to_planar(qbc)).cuda()
    out = torch.empty_like(qin)
    cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
    K.step2_split(rp, qin, out, cfl, [4, 4], [4, 4], [1, 2], meth, [4] * 3, dx, dy, dt,
                  gin=L.geom2(mx, my, ng), num_ghost=ng, aux=_dev(abc),
                  aux_bc_lower=[4, 4], aux_bc_upper=[4, 4], whole=True)
    torch.cuda.synchronize()
    got = from_planar(out.cpu().numpy())
    assert np.array_equal(got, ref), np.abs(got - ref).max()


@pytest.mark.parametrize("kind", [O.ACOUSTICS, O.EULER])
def test_fwave_flag_steps_bit_exact(kind):
    """use_fwave with the acoustics and Euler solvers' waves taken as f-waves:
    unsplit (transverse 1 and 2, where trans_cor folds the f-wave correction
    into the transverse split) and split, bitwise vs the oracle."""
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device
    from tests.test_gpu_classic2 import KIND_PD, _oracle_rp
    print("f-wave flag", kind, flush=True)
    rng = np.random.default_rng(50 + kind)
    mx, my, ng = 61, 38, 2
    mw = O.MWAVES[kind]
    bcl, bcu = [3, 1], [1, 3]
    for trans, split in ((1, False), (2, False), (0, True)):
        q = seeded_state(rng, kind, (mx, my))
        dx, dy, dt = 1.0 / mx, 0.8 / my, 0.2 / mx
        qbc = O.with_ghosts(q, ng)
        O.bc2(qbc, ng, bcl, bcu)
        qold = qbc.copy(order="F")
        if split:
            meth = O.method_array(dimensional_split=True)
            c = max(O.step2(_oracle_rp(kind), qold, qbc, ng, dx, dy, dt, meth, [4] * mw, ids=1,
                            use_fwave=True),
                    O.step2(_oracle_rp(kind), qbc, qbc, ng, dx, dy, dt, meth, [4] * mw, ids=2,
                            use_fwave=True))
        else:
            meth = O.method_array(transverse=trans)
            c = O.step2(_oracle_rp(kind), qold, qbc, ng, dx, dy, dt, meth, [4] * mw,
                        use_fwave=True)
        ref = O.interior(qbc, ng)
        rp = K.make_rp(kind, KIND_PD[kind], fwave=True)
        out = torch.empty_like(_dev(q))
        cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
        if split:
            K.step2_split(rp, _dev(q), out, cfl, bcl, bcu, [1, 2], meth, [4] * mw, dx, dy, dt)
        else:
            K.step2_unsplit(rp, _dev(q), out, cfl, bcl, bcu, [1, 2], meth, [4] * mw, dx, dy, dt)
        torch.cuda.synchronize()
        got = q_from_device(out)
        assert np.array_equal(got, ref), (trans, split, np.abs(got - ref).max())
        assert float(cfl.item()) == c


def test_fmod_step2ds_fwave():
    """The f2py drop-in: fmod.step2ds with use_fwave and the f-wave solver's tag."""
    from pyclaw_amd import riemann
    from pyclaw_amd.fmod import classic2
    print("fmod step2ds f-wave", flush=True)
    rng = np.random.default_rng(47)
    mx, my, ng = 30, 22, 2
    q, aux = _swe_bathymetry(rng, mx, my)
    qbc, abc = O.with_ghosts(q, ng), O.with_ghosts(aux, ng)
    O.bc2(qbc, ng, [1, 3], [3, 1])
    O.bc2(abc, ng, [1, 1], [1, 1], reflect=(-1, -1))
    meth = O.method_array(dimensional_split=True, num_aux=1)
    dx, dy, dt = 1.0 / mx, 1.0 / my, 0.003
    rp_o = O.make_rp(O.SHALLOW_FWAVE, grav=1.0)
    classic2.cparam.grav = 1.0
    classic2.cparam.dry_tolerance = 1e-3
    for ids in (1, 2):
        ref = qbc.copy(order="F")
        c_ref = O.step2(rp_o, qbc, ref, ng, dx, dy, dt, meth, [1] * 3, ids=ids, use_fwave=True,
                        aux=abc)
        out = qbc.copy(order="F")
        rpn = riemann.shallow_bathymetry_fwave_2D.rpn2._cpointer
        out, c = classic2.step2ds(max(mx, my), ng, mx, my, qbc, out, abc, dx, dy, dt, meth,
                                  [1] * 3, None, None, None, None, ids, True, rpn, rpn)
        assert np.array_equal(out, ref) and c == c_ref, ids


def _sill_claw(precision="fp64"):
    """examples/shallow_2d/sill.py:41-86 through the API (output off)."""
    import pyclaw_amd as pyclaw
    from pyclaw_amd import riemann
    _, _, _, wave_maker = C.sill()

    def wave_maker_bc(state, dim, t, qbc, auxbc, num_ghost):
        if dim.on_lower_boundary:
            wave_maker(state.t, qbc, num_ghost)

    solver = pyclaw.ClawSolver2D(riemann.shallow_bathymetry_fwave_2D)
    solver.dimensional_split = 1
    solver.bc_lower[0] = pyclaw.BC.custom
    solver.user_bc_lower = wave_maker_bc
    solver.bc_upper[0] = pyclaw.BC.extrap
    solver.bc_lower[1] = pyclaw.BC.wall
    solver.bc_upper[1] = pyclaw.BC.wall
    for k in range(2):
        solver.aux_bc_lower[k] = pyclaw.BC.extrap
        solver.aux_bc_upper[k] = pyclaw.BC.extrap
    solver.precision = precision
    x = pyclaw.Dimension(0., 4., 80, name="x")
    y = pyclaw.Dimension(0., 1., 20, name="y")
    domain = pyclaw.Domain([x, y])
    state = pyclaw.State(domain, 3, num_aux=1)
    X, Y = state.p_centers
    state.aux[0, :, :] = 0.8 * np.exp(-10 * ((X - 1.) ** 2 + (Y - 0.5) ** 2))
    state.q[0, :, :] = 1. - state.aux[0, :, :]
    state.q[1, :, :] = 0.
    state.q[2, :, :] = 0.
    state.problem_data["grav"] = 1.0
    state.problem_data["dry_tolerance"] = 1.e-3
    state.problem_data["sea_level"] = 0.
    claw = pyclaw.Controller()
    claw.tfinal = 10
    claw.solution = pyclaw.Solution(state, domain)
    claw.solver = solver
    claw.num_output_times = 40
    claw.keep_copy = True
    claw.output_format = None
    return claw


def test_api_sill_fixture(golden_dir):
    """examples/shallow_2d/test_shallow2d.py:38-39: depth at t = 10 within L1
    1e-6 of the fixture; the run equals the oracle's run bit for bit (dt
    sequence and every frame), which the persistent ghost cells of the wave
    maker's qbc require."""
    from tests.test_oracle import sill_run
    print("sill through the API", flush=True)
    claw = _sill_claw()
    claw.run()
    run, frames, d = sill_run()
    assert claw.solver.status["numsteps"] == run.numsteps
    for k in (1, 20, 40):
        assert np.array_equal(claw.frames[k].state.q, frames[k]), k
    exp = np.load(f"{golden_dir}/shallow_2d_expected.npz")["sill"]
    depth = claw.frames[40].state.q[0].reshape(-1)
    err = d[0] * d[1] * np.abs(depth - exp).sum()
    print("sill L1", err)
    assert err < 1e-6


def test_api_sill_fp32k_close():
    """The sill run in FP32 + Kahan: relative max of the depth within 1e-5 of
    FP64 at t = 10 (about 270 steps)."""
    print("sill fp32k", flush=True)
    a, b = _sill_claw("fp64"), _sill_claw("fp32k")
    a.run()
    b.run()
    qa, qb = a.frames[40].state.q, b.frames[40].state.q
    err = np.abs(qa[0] - qb[0]).max() / np.abs(qa[0]).max()
    print("sill fp32k rel max", err)
    assert err < 1e-5
