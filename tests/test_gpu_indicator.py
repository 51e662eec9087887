"""GPU parity of the learned troubled-interface indicator (indicator.hip)
against its CPU restatement (oracle/mlp.py), of the flag-selected SharpClaw
reconstruction, and the accuracy/throughput gain of the FP32 + indicator path
over FP64 full WENO5."""
import numpy as np
import pytest
import torch

from oracle import mlp as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _swe_state(rng, mx, my):
    """Smooth shallow-water state with seeded phases plus two jumps (one along
    x, one along y): some interfaces troubled, most not."""
    x = (np.arange(mx) + 0.5) / mx
    y = (np.arange(my) + 0.5) / my
    X, Y = np.meshgrid(x, y, indexing="ij")
    a, b, c = rng.uniform(0, 2 * np.pi, 3)
    h = 1.0 + 0.2 * np.sin(2 * np.pi * X + a) * np.cos(2 * np.pi * Y + b)
    h = h + 0.5 * (X > 0.37) + 0.3 * (Y > 0.61)
    q = np.zeros((3, mx, my), order="F")
    q[0] = h
    q[1] = h * 0.2 * np.cos(2 * np.pi * Y + c)
    q[2] = h * 0.1 * np.sin(2 * np.pi * X + c)
    return q


def _dev(q):
    from pyclaw_amd.device import q_to_device
    return q_to_device(q)


def _planar(q):
    return np.ascontiguousarray(np.transpose(q, (0, 2, 1)))


def _params(seed):
    p = M.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    p[M.P_W1 + 6 * M.NH:M.P_W1 + 7 * M.NH] = rng.normal(0, 0.3, M.NH)
    p[M.P_B2:M.P_W3] = rng.normal(0, 0.3, M.NH)
    p[M.P_B3] = 0.1
    return p


@pytest.mark.parametrize("bcl,bcu", [([1, 1], [1, 1]), ([2, 3], [2, 1]), ([3, 2], [3, 2])])
def test_features_labels_match_oracle(bcl, bcu):
    from pyclaw_amd import indicator as I
    rng = np.random.default_rng(sum(bcl) * 10 + sum(bcu))
    mx, my = 37, 29
    q = _swe_state(rng, mx, my)
    flat, wtol = 1e-6, 0.05
    f_ref, l_ref = M.features(_planar(q), bcl, bcu, flat, wtol)
    f, l = I.features(_dev(q), bcl, bcu, flat, wtol)
    assert np.array_equal(f.cpu().numpy(), f_ref)
    assert np.array_equal(l.cpu().numpy(), l_ref)
    assert 0 < l_ref.sum() < l_ref.size


def test_logits_match_oracle():
    # checks the MFMA operand/result lane layout of every layer
    from pyclaw_amd import indicator as I
    rng = np.random.default_rng(1)
    feat = np.zeros((1000, 8), np.float32)       # not a multiple of 16: padded tile
    feat[:, :6] = rng.uniform(-1, 1, (1000, 6))
    feat[:, 6] = 1.0
    p = _params(2)
    z_ref, _ = M.forward(p, feat)
    z = I.logits(torch.tensor(feat, device="cuda"),
                 torch.tensor(p, device="cuda")).cpu().numpy()
    assert np.abs(z - z_ref).max() <= 2e-3 * (1 + np.abs(z_ref).max())
    # a transposed or shifted layout would decorrelate the two completely
    assert np.corrcoef(z, z_ref)[0, 1] > 0.9999


def _train_call(feat, label, idx, p, step, lr=1e-3, pw=3.0):
    from pyclaw_amd import kernels as K
    dev = "cuda"
    P = torch.tensor(p, device=dev)
    G = torch.zeros_like(P)
    m1 = torch.zeros_like(P)
    m2 = torch.zeros_like(P)
    loss = torch.zeros(1, device=dev)
    K.mlp_train_step(torch.tensor(feat, device=dev), torch.tensor(label, device=dev),
                     torch.tensor(idx, dtype=torch.int32, device=dev), P, G, m1, m2, step, lr,
                     pw, loss)
    torch.cuda.synchronize()
    return P.cpu().numpy(), G.cpu().numpy(), float(loss.item())


def test_gradients_match_oracle():
    rng = np.random.default_rng(3)
    n = 5000
    feat = np.zeros((n, 8), np.float32)
    feat[:, :6] = rng.uniform(-1, 1, (n, 6))
    feat[:, 6] = 1.0
    label = (rng.uniform(size=n) < 0.3).astype(np.uint8)
    idx = rng.integers(0, n, 1000)
    idx[:3] = [-1, n, n + 7]                      # out-of-range entries are skipped
    p = _params(4)
    _, g, loss = _train_call(feat, label, idx, p, 0)
    valid = idx[(idx >= 0) & (idx < n)]
    # the kernel averages over nbatch (all entries); the oracle over its batch
    lref, gref = M.loss_grad(p, feat, label, valid, 3.0)
    scale = len(valid) / len(idx)
    assert abs(loss - lref * scale) < 1e-4 * abs(lref)
    groups = {"W1": (M.P_W1, M.P_W2), "W2": (M.P_W2, M.P_B2), "b2": (M.P_B2, M.P_W3),
              "w3": (M.P_W3, M.P_B3), "b3": (M.P_B3, M.P_N)}
    for name, (a, b) in groups.items():
        ga, gr = g[a:b], gref[a:b] * scale
        err = np.linalg.norm(ga - gr) / max(np.linalg.norm(gr), 1e-30)
        assert err < 1e-2, (name, err)
    assert np.all(g[M.P_W1 + 7 * M.NH:M.P_W2] == 0)


def test_adam_step_matches_torch():
    rng = np.random.default_rng(5)
    n = 16                                         # one tile: one wavefront
    feat = np.zeros((n, 8), np.float32)
    feat[:, :6] = rng.uniform(-1, 1, (n, 6))
    feat[:, 6] = 1.0
    label = (np.arange(n) % 3 == 0).astype(np.uint8)
    idx = np.arange(n)
    p = _params(6)
    _, g, _ = _train_call(feat, label, idx, p, 0)
    p1, g1, _ = _train_call(feat, label, idx, p, 1, lr=1e-3)
    assert np.all(g1 == 0)                         # Adam clears the gradient
    t = torch.tensor(p.astype(np.float64), requires_grad=True)
    opt = torch.optim.Adam([t], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    t.grad = torch.tensor(g.astype(np.float64))
    opt.step()
    ref = t.detach().numpy()
    ref[M.P_W1 + 7 * M.NH:M.P_W2] = 0.0
    assert np.abs(p1 - ref).max() < 2e-5


@pytest.mark.parametrize("meqn", [1, 3, 4, 5])
def test_flags_match_oracle(meqn):
    """meqn <= 4 runs the one-interface-per-lane kernel, larger meqn the
    (interface, component)-per-lane kernel; both against the oracle."""
    from pyclaw_amd import indicator as I
    rng = np.random.default_rng(8)
    mx, my = 33, 26
    q = _swe_state(rng, mx, my)
    if meqn < 3:
        q = np.asfortranarray(q[:meqn])
    elif meqn > 3:
        extra = [q[0] * (1.0 + 0.1 * k) + 0.2 * k * np.roll(q[1], k, axis=0)
                 for k in range(1, meqn - 2)]
        q = np.asfortranarray(np.concatenate([q] + [e[None] for e in extra], axis=0))
    p = _params(9)
    ind = I.TroubledInterfaceIndicator(p, threshold=0.2, flat=1e-6)
    fl = ind.flags(_dev(q), [1, 2], [3, 2]).cpu().numpy()
    ref, margin = M.flags(_planar(q), [1, 2], [3, 2], 1e-6, p, 0.2)
    sure = margin > 1e-2
    assert sure.mean() > 0.9
    assert np.array_equal(fl[sure], ref[sure])
    assert 0 < ref.sum() < ref.size


def _advection_dq(q, flags, dt, dx, dy, u, v):
    """dt*L(q) of periodic advection with u, v > 0 where each interface uses
    the linear WENO5 weights unless it or a normal neighbour is flagged."""
    mx, my = q.shape[1], q.shape[2]
    ng = 3
    qbc = O.with_ghosts(q, ng)
    O.bc2(qbc, ng, [2, 2], [2, 2])
    nxi = (mx + 1) * my
    fx = flags[:nxi].reshape(my, mx + 1)
    fy = flags[nxi:].reshape(my + 1, mx)

    def lin(st):   # right edge of cell 2 of the stencil st[0..4]
        return (2 * st[0] - 13 * st[1] + 47 * st[2] + 27 * st[3] - 3 * st[4]) / 60.0

    dq = np.zeros((1, mx, my))
    # x: upwind edge of interface c (1..mx+1) is the right edge of cell c-1
    for j in range(my):
        row = qbc[:, :, j + ng]
        _, qr = O.weno5(row, ng)
        QR = np.zeros(mx + 2)
        for c in range(1, mx + 2):
            f = fx[j, c - 1] | fx[j, max(c - 2, 0)] | fx[j, min(c, mx)]
            a = c - 2 + ng                          # 0-based array index of cell c-1
            QR[c] = qr[0, a] if f else lin(row[0, a - 2:a + 3])
        dq[0, :, j] -= dt / dx * u * (QR[2:mx + 2] - QR[1:mx + 1])
    for i in range(mx):
        col = qbc[:, i + ng, :]
        _, qt = O.weno5(col, ng)
        QT = np.zeros(my + 2)
        for r in range(1, my + 2):
            f = fy[r - 1, i] | fy[max(r - 2, 0), i] | fy[min(r, my), i]
            a = r - 2 + ng
            QT[r] = qt[0, a] if f else lin(col[0, a - 2:a + 3])
        dq[0, i, :] -= dt / dy * v * (QT[2:my + 2] - QT[1:my + 1])
    return dq


def _dilated_ids(fl, mx, my):
    """Ids of the interfaces a flagged stage reconstructs with full WENO5: an
    x-interface whose flag or a row neighbour's is set, a y-interface whose
    flag or a column neighbour's is set (neighbours clamped at the ends)."""
    nx = (mx + 1) * my
    fx = fl[:nx].reshape(my, mx + 1).astype(bool)
    fy = fl[nx:].reshape(my + 1, mx).astype(bool)
    dx = fx.copy()
    dx[:, 1:] |= fx[:, :-1]
    dx[:, :-1] |= fx[:, 1:]
    dy = fy.copy()
    dy[1:] |= fy[:-1]
    dy[:-1] |= fy[1:]
    return np.concatenate([np.flatnonzero(dx.ravel()), nx + np.flatnonzero(dy.ravel())])


@pytest.mark.parametrize("mx,my,frac", [(3, 3, 0.5), (4, 5, 0.3), (255, 3, 0.05),
                                        (256, 7, 0.5), (1000, 9, 0.01), (9000, 3, 0.02),
                                        (300, 301, 1.0), (64, 64, 0.0)])
def test_flag_list_matches_dilated_flags(mx, my, frac):
    """clawhip_mlp_flag_list: the listed ids are exactly the dilated flags, each
    once (the order is free), on the smallest grid the entry point takes (3 x 3),
    rows narrower than a wavefront, rows split over several blocks, all and no
    flags; a second call over the same buffers starts
    its count again from zero."""
    from pyclaw_amd import kernels as K
    print("flag list", mx, my, frac, flush=True)
    rng = np.random.default_rng(mx * 7 + my)
    n = K.mlp_num_interfaces(mx, my)
    out = None
    for rep in range(2):
        fl = (rng.uniform(size=n) < frac).astype(np.uint8)
        out = K.mlp_flag_list(torch.tensor(fl, device="cuda"), mx, my, out)
        cnt = int(out[1].item())
        got = np.sort(out[0][:cnt].cpu().numpy())
        want = _dilated_ids(fl, mx, my)
        assert cnt == want.size, (rep, cnt, want.size)
        assert np.array_equal(got, want), rep


@pytest.mark.parametrize("pattern", ["none", "ones", "zeros", "random"])
def test_stage_flag_selection(pattern):
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device
    mx, my = 41, 35
    rng = np.random.default_rng(11)
    x = (np.arange(mx) + 0.5) / mx
    y = (np.arange(my) + 0.5) / my
    X, Y = np.meshgrid(x, y, indexing="ij")
    q = np.asfortranarray((np.sin(2 * np.pi * X) * np.cos(4 * np.pi * Y) + (X > 0.5))[None])
    n = K.mlp_num_interfaces(mx, my)
    fl = {"none": None, "ones": np.ones(n, np.uint8), "zeros": np.zeros(n, np.uint8),
          "random": (rng.uniform(size=n) < 0.2).astype(np.uint8)}[pattern]
    u, v = 0.5, 0.7
    pd = dict(u=u, v=v)
    dx, dy, dt = 1.0 / mx, 1.0 / my, 0.3 / max(mx, my)
    qi = _dev(q)
    out = torch.empty_like(qi)
    cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
    fdev = None if fl is None else torch.tensor(fl, device="cuda")
    K.sharp2_stage(K.make_rp(O.ADVECTION, pd), 0, out, cfl, [2, 2], [2, 2], [1, 2], dt, dx, dy,
                   qi, flags=fdev)
    got = q_from_device(out)
    if pattern in ("none", "ones"):
        rp = O.make_rp(O.ADVECTION, u=u, v=v)
        qbc = O.with_ghosts(q, 3)
        O.bc2(qbc, 3, [2, 2], [2, 2])
        ref = O.interior(O.sharp_flux2(rp, qbc, 3, dx, dy, dt, 0)[0], 3)
        assert np.array_equal(got, ref)
    else:
        ref = _advection_dq(q, fl, dt, dx, dy, u, v)
        assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max()


def test_stage_all_flags_equals_char_weno():
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device
    rng = np.random.default_rng(12)
    mx, my = 50, 21
    q = _swe_state(rng, mx, my)
    qi = _dev(q)
    outs = []
    for fl in (None, torch.ones(K.mlp_num_interfaces(mx, my), dtype=torch.uint8, device="cuda")):
        out = torch.empty_like(qi)
        cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
        K.sharp2_stage(K.make_rp(O.SHALLOW, dict(grav=1.0)), 2, out, cfl, [1, 3], [3, 1],
                       [1, 2], 1e-3, 1.0 / mx, 1.0 / my, qi, flags=fl)
        outs.append(q_from_device(out))
    assert np.array_equal(outs[0], outs[1])



def _stage_ms(K, rp, q, flags=None, flag_list=None, recon=None, reps=5, n=10):
    out = torch.empty_like(q)
    cfl = torch.zeros(1, dtype=torch.float64, device=q.device)
    my, mx = q.shape[1], q.shape[2]
    args = (rp, 2, out, cfl, [2, 2], [2, 2], [1, 2], 1e-4, 1.0 / mx, 1.0 / my, q)
    kw = dict(flags=flags, flag_list=flag_list, recon=recon)
    K.sharp2_stage(*args, **kw)
    return _best_ms(lambda: K.sharp2_stage(*args, **kw), reps, n)


def _best_ms(fn, reps=5, n=10):
    best = None
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(n):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / n
        best = ms if best is None else min(best, ms)
    return best


def test_indicator_gain_accuracy_and_speed():
    """The shipped network on the periodic shallow-water family (humps + dam
    breaks, characteristic WENO5, SSP33, CFL 0.45 / max 0.5): FP32 + Kahan with
    the learned selection (flags refreshed every second step) stays within 1e-4
    relative L1 of FP64 full WENO5 over 100 steps at 256^2 while most interfaces
    take the linear weights; at 2048^2 its step (three flagged FP32 stages plus
    half a flags pass) is at least 1.8x faster than the FP64
    full-WENO step (three stages), and faster than the FP32 full-WENO step."""
    from pyclaw_amd import indicator as I
    from pyclaw_amd import kernels as K
    res = {}
    for name, prec, ind in (("fp64", "fp64", None), ("mlp", "fp32k", True)):
        indicator = I.TroubledInterfaceIndicator() if ind else None
        solver = I.swe_solver(prec, indicator)
        sol = I.swe_solution(3, 256, 256)
        solver.dt = 0.45 / 256 / 1.6
        solver.setup(sol)
        frac = []
        for k in range(100):
            solver.evolve_to_time(sol)
            if indicator is not None and k % 10 == 0:
                frac.append(indicator.troubled_fraction())
        if indicator is not None:
            assert indicator.refresh_interval(solver.cfl_max) == 2
        res[name] = (sol.state.q.copy(), sol.t, np.mean(frac) if frac else None)
    qa, ta, _ = res["fp64"]
    qb, tb, frac = res["mlp"]
    err = np.abs(qb - qa).sum() / np.abs(qa).sum()
    print(f"256^2 x 100 steps: rel L1 {err:.2e}, troubled interfaces {frac:.3f}")
    # the two runs choose their own dt from their own CFL numbers
    assert abs(tb - ta) < 1e-4 * ta
    assert err < 1e-4
    assert frac < 0.5
    # kernel-level step cost at 2048^2 on an evolved state
    n = 2048
    solver = I.swe_solver("fp64")
    sol = I.swe_solution(3, n, n)
    solver.dt = 0.45 / n / 1.6
    solver.setup(sol)
    for _ in range(10):
        solver.evolve_to_time(sol)
    q64 = sol.state.device_q(K.L.F64)[0].contiguous()
    q32 = q64.float()
    rp = K.make_rp(3, {"grav": 1.0, "efix": 1})
    ind = I.TroubledInterfaceIndicator()
    fl = ind.flags(q32, [2, 2], [2, 2])
    torch.cuda.synchronize()
    t64 = _stage_ms(K, rp, q64)
    t32 = _stage_ms(K, rp, q32)
    t32f = _stage_ms(K, rp, q32, fl, ind.flag_list, ind.recon_buffer(q32))
    tflags = _best_ms(lambda: ind.flags(q32, [2, 2], [2, 2]))
    step64 = 3 * t64
    step32 = 3 * t32
    step_mlp = 3 * t32f + tflags / 2
    print(f"2048^2 stage ms: fp64 {t64:.3f} fp32 {t32:.3f} fp32 flagged {t32f:.3f} "
          f"flags {tflags:.3f} troubled {ind.troubled_fraction():.4f}; step ratio "
          f"fp64/mlp {step64 / step_mlp:.2f}")
    assert step_mlp < step32
    assert step64 / step_mlp >= 1.8
