"""Thin Python wrappers over the C ABI: validate shapes, pass device pointers.

Tensors are torch tensors on the HIP device in planar layout (see device.py).
All calls are asynchronous on torch's current stream.
"""
import ctypes

import torch

from . import _lib as L


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _dtype_code(t, fast=False):
    """The entry point's dtype for an array of t's type: FP64 (fast: the
    FMA-contracted FP64 variant, CLAWHIP_F64_FAST) or FP32 + Kahan."""
    if t.dtype == torch.float64:
        return L.F64_FAST if fast else L.F64
    if t.dtype == torch.float32:
        return L.F32K
    raise L.ClawhipError(f"unsupported dtype {t.dtype}")


def make_rp(kind, problem_data, efix=1, meqn=0, fwave=False):
    """Parameters of the Riemann solver from problem_data (state.py:206-231);
    meqn: the number of fields, for solvers that take any (vc advection);
    fwave: the solver returns f-waves (solver.fwave, step2's use_fwave)."""
    pd = problem_data
    rho = float(pd.get("rho", 1.0))
    bulk = float(pd.get("bulk", 1.0))
    cc = float(pd.get("cc", (bulk / rho) ** 0.5))
    zz = float(pd.get("zz", rho * cc))
    return L.RP(kind, int(pd.get("efix", efix)), rho, bulk, cc, zz,
                float(pd.get("gamma", 1.4)), float(pd.get("grav", 1.0)),
                float(pd.get("u", 0.0)), float(pd.get("v", 0.0)), int(meqn), int(bool(fwave)),
                float(pd.get("dry_tolerance", 1e-3)), float(pd.get("sea_level", 0.0)))


def cfl_reset(cfl):
    L.check(L.lib().clawhip_cfl_reset(_ptr(cfl), _stream()), "clawhip_cfl_reset")


_RP_MEQN = {1: 3, 2: 4, 3: 3, 4: 1, 5: 5, 7: 1, 8: 3, 9: 3, 10: 5, 11: 3}


def _check_fields(rp, meqn, aux, method):
    """The kernels read as many fields as the solver has (kind 6: rp.meqn) and
    method[6] aux fields: the tensors must hold exactly those."""
    want = rp.meqn if rp.kind == 6 else _RP_MEQN.get(rp.kind)
    if want is not None and meqn != want:
        raise ValueError("q has %d fields; the Riemann solver takes %d" % (meqn, want))
    if aux is not None and aux.shape[0] != int(method[6]):
        raise ValueError("aux has %d fields, method[6] says %d" % (aux.shape[0], int(method[6])))


def _aux_geom(aux, mx, my):
    """Geometry of a planar aux tensor: compact (maux, my, mx) or with stored
    ghosts (maux, my + 2g, mx + 2g)."""
    if aux is None:
        return None
    g = (aux.shape[2] - mx) // 2
    if aux.shape[1] != my + 2 * g or aux.shape[2] != mx + 2 * g or g < 0:
        raise ValueError("aux shape %s does not fit a %d x %d grid" % (tuple(aux.shape), mx, my))
    return L.geom2(mx, my, g)


def step2_unsplit(rp, q_in, q_out, cfl, bc_lower, bc_upper, reflect, method, mthlim,
                  dx, dy, dt, comp_in=None, comp_out=None, num_ghost=2, gin=None, aux=None,
                  aux_bc_lower=None, aux_bc_upper=None, fast=False):
    """One unsplit step q_in -> q_out (planar (meqn, my, mx) tensors).
    aux: planar aux fields of q's dtype, compact or with num_ghost stored ghosts
    (then with aux codes 4 = given), for solvers that read aux; method[5] > 0
    names a capacity field among them."""
    meqn, my, mx = q_out.shape
    gout = L.geom2(mx, my)
    if gin is None:
        gin = gout
    code = _dtype_code(q_out, fast)
    assert q_in.dtype == q_out.dtype and q_in.is_contiguous() and q_out.is_contiguous()
    if aux is not None:
        assert aux.dtype == q_out.dtype and aux.is_contiguous()
    _check_fields(rp, meqn, aux, method)
    gaux = _aux_geom(aux, mx, my)
    rc = L.lib().clawhip_step2_unsplit(
        ctypes.byref(rp), code, _ptr(q_in), ctypes.byref(gin), _ptr(q_out), ctypes.byref(gout),
        _ptr(comp_in), _ptr(comp_out), L.ints(bc_lower), L.ints(bc_upper),
        L.ints(reflect), int(num_ghost), L.ints(method),
        L.ints(list(mthlim) + [0] * (4 - len(mthlim))), float(dx), float(dy), float(dt),
        _ptr(aux), ctypes.byref(gaux) if aux is not None else None,
        L.ints(aux_bc_lower) if aux is not None else None,
        L.ints(aux_bc_upper) if aux is not None else None, _ptr(cfl), _stream())
    L.check(rc, "clawhip_step2_unsplit")


def _work_like(q_out, ng, work):
    """Padded intermediate (meqn, [mz+2ng,] my+2ng, mx+2ng), reused between calls."""
    shape = (q_out.shape[0],) + tuple(n + 2 * ng for n in q_out.shape[1:])
    if work is None or tuple(work.shape) != shape or work.dtype != q_out.dtype:
        work = torch.empty(shape, dtype=q_out.dtype, device=q_out.device)
    return work


def step2_split(rp, q_in, q_out, cfl, bc_lower, bc_upper, reflect, method, mthlim, dx, dy, dt,
                comp_in=None, comp_out=None, num_ghost=2, gin=None, work=None, aux=None,
                aux_bc_lower=None, aux_bc_upper=None, whole=False, fast=False):
    """Both dimension-split sweeps (step2ds ids=1, 2); returns the work buffers.
    aux: planar (maux, my, mx) tensor of q's dtype for solvers with aux fields.
    whole: q_in and q_out are whole qbc arrays (num_ghost stored ghosts); q_out
    receives every cell the reference's sweeps write (the sweeps run in it)."""
    meqn, my, mx = q_out.shape
    gout = L.geom2(mx, my)
    if whole:
        my, mx = my - 2 * num_ghost, mx - 2 * num_ghost
        gout = L.geom2(mx, my, num_ghost)
    if gin is None:
        gin = gout
    code = _dtype_code(q_out, fast)
    _check_fields(rp, meqn, aux, method)
    w, cw = work if work is not None else (None, None)
    if not whole:
        w = _work_like(q_out, num_ghost, w)
        if code == L.F32K:
            cw = _work_like(q_out, num_ghost, cw)
    rc = L.lib().clawhip_step2_split(
        ctypes.byref(rp), code, _ptr(q_in), ctypes.byref(gin), _ptr(q_out), ctypes.byref(gout),
        _ptr(comp_in), _ptr(comp_out), _ptr(w), _ptr(cw), L.ints(bc_lower), L.ints(bc_upper),
        L.ints(reflect), int(num_ghost), L.ints(method),
        L.ints(list(mthlim) + [0] * (4 - len(mthlim))), float(dx), float(dy), float(dt),
        _ptr(aux), ctypes.byref(_aux_geom(aux, mx, my)) if aux is not None else None,
        L.ints(aux_bc_lower) if aux is not None else None,
        L.ints(aux_bc_upper) if aux is not None else None, _ptr(cfl), _stream())
    L.check(rc, "clawhip_step2_split")
    return (w, cw)


def step3_unsplit(rp, q_in, q_out, cfl, bc_lower, bc_upper, reflect, method, mthlim, dx, dy,
                  dz, dt, num_ghost=2, gin=None, work=None):
    """The unsplit 3D step (step3), FP64: q_in (compact, or with stored ghosts
    given by gin) -> q_out (planar (meqn, mz, my, mx)); returns the work buffer."""
    meqn, mz, my, mx = q_out.shape
    gout = L.geom3(mx, my, mz)
    if gin is None:
        gin = gout
    n = int(L.lib().clawhip_step3_unsplit_work_elems(meqn, mx, my, mz))
    if work is None or work.numel() != n:
        work = torch.empty(n, dtype=torch.float64, device=q_out.device)
    rc = L.lib().clawhip_step3_unsplit(
        ctypes.byref(rp), _dtype_code(q_out), _ptr(q_in), ctypes.byref(gin), _ptr(q_out),
        ctypes.byref(gout), _ptr(work), L.ints(bc_lower), L.ints(bc_upper), L.ints(reflect),
        int(num_ghost), L.ints(method), L.ints(list(mthlim) + [0] * (4 - len(mthlim))),
        float(dx), float(dy), float(dz), float(dt), _ptr(cfl), _stream())
    L.check(rc, "clawhip_step3_unsplit")
    return work


def step3_split(rp, q_in, q_out, cfl, bc_lower, bc_upper, reflect, method, mthlim, dx, dy, dz,
                dt, comp_in=None, comp_out=None, num_ghost=2, work=None, fast=False):
    """The three dimension-split sweeps of ClawSolver3D (step3ds idir=1,2,3)."""
    meqn, mz, my, mx = q_out.shape
    g = L.geom3(mx, my, mz)
    code = _dtype_code(q_out, fast)
    w, cw = work if work is not None else (None, None)
    w = _work_like(q_out, num_ghost, w)
    if code == L.F32K:
        cw = _work_like(q_out, num_ghost, cw)
    rc = L.lib().clawhip_step3_split(
        ctypes.byref(rp), code, _ptr(q_in), ctypes.byref(g), _ptr(q_out), ctypes.byref(g),
        _ptr(comp_in), _ptr(comp_out), _ptr(w), _ptr(cw), L.ints(bc_lower), L.ints(bc_upper),
        L.ints(reflect), int(num_ghost), L.ints(method),
        L.ints(list(mthlim) + [0] * (4 - len(mthlim))), float(dx), float(dy), float(dz),
        float(dt), _ptr(cfl), _stream())
    L.check(rc, "clawhip_step3_split")
    return (w, cw)


# SharpClaw stage modes (include/clawhip.h, clawhip_sharp2_stage)
Y_ARRAY, Y_AXPY, Y_AXPY6 = 0, 1, 2
OUT_DQDT, OUT_SSP33_2, OUT_SSP33_3, OUT_ADD6, OUT_SSP104_FIN, OUT_F32_S2, OUT_F32_S3, OUT_D, \
    OUT_SSP22 = range(9)


def sharp2_stage(rp, char_decomp, out, cfl, bc_lower, bc_upper, reflect, dt, dx, dy, P,
                 Q=None, s=0.0, y_mode=Y_ARRAY, out_mode=OUT_D, A=None, B=None, aux_in=None,
                 comp_in=None, aux_out=None, comp_out=None, flags=None, flag_list=None,
                 recon=None, qaux=None, qaux_bc_lower=None, qaux_bc_upper=None, src=None,
                 fast=False):
    """D = dt*L(Y) (SharpClaw flux2) fused with one RK stage; planar (meqn, my, mx).
    flags: uint8 troubled-interface flags (mlp_flags) selecting full WENO5 per
    interface (the linear weights elsewhere), with flag_list = (ids, count) from
    mlp_flag_list and recon a work tensor of recon_elems(...) elements of the
    state's dtype (both made here when None); flags None = full WENO5
    everywhere.  qaux: the solver's aux fields (planar, compact or with 3 stored
    ghosts and codes 4) for the solvers that read them.  P may carry 3 stored
    ghosts (planar (meqn, my+6, mx+6), boundary codes 4), filled by the caller;
    src: planar interior dq_src increment."""
    meqn, my, mx = out.shape
    if flags is not None:
        if flag_list is None:
            flag_list = mlp_flag_list(flags, mx, my)
        if flag_list[0].numel() < mlp_num_interfaces(mx, my):
            raise ValueError("flag_list ids must hold (mx+1)*my + mx*(my+1) entries")
        if recon is None:
            recon = torch.empty(recon_elems(mx, my, meqn), dtype=out.dtype, device=out.device)
        assert recon.dtype == out.dtype and recon.numel() >= recon_elems(mx, my, meqn)
    _check_fields(rp, meqn, None, None)
    for t in (P, Q, A, B, src):
        assert t is None or (t.dtype == out.dtype and t.numel() >= out.numel())
    assert src is None or src.shape == out.shape
    g = L.geom2(mx, my)
    gp = (P.shape[2] - mx) // 2
    assert P.shape[1] == my + 2 * gp and P.shape[2] == mx + 2 * gp and P.shape[0] == meqn
    gin = L.geom2(mx, my, gp) if gp else None
    code = _dtype_code(out, fast)
    gaux = _aux_geom(qaux, mx, my)
    if qaux is not None:
        assert qaux.dtype == out.dtype and qaux.is_contiguous()
    rc = L.lib().clawhip_sharp2_stage(
        ctypes.byref(rp), code, int(char_decomp), ctypes.byref(g), L.ints(bc_lower),
        L.ints(bc_upper), L.ints(reflect), int(y_mode), int(out_mode), float(dt), float(s),
        float(dx), float(dy), _ptr(P), ctypes.byref(gin) if gin is not None else None, _ptr(Q),
        _ptr(A), _ptr(B), _ptr(aux_in), _ptr(comp_in), _ptr(out), _ptr(aux_out), _ptr(comp_out),
        _ptr(src), _ptr(flags), _ptr(flag_list[0]) if flags is not None else None,
        _ptr(flag_list[1]) if flags is not None else None, _ptr(recon), _ptr(qaux),
        ctypes.byref(gaux) if qaux is not None else None,
        L.ints(qaux_bc_lower) if qaux is not None else None,
        L.ints(qaux_bc_upper) if qaux is not None else None,
        int(qaux.shape[0]) if qaux is not None else 0, _ptr(cfl), _stream())
    L.check(rc, "clawhip_sharp2_stage")


def sharp2_ssp33_run(rp, char_decomp, bufs, comps, ctl_dev, bc_lower, bc_upper, reflect, dx, dy,
                     dq_dt, r2, aux23, nsteps, indicator=None, dq_dt_full=None, recon=None,
                     positive=(), stage_events=None, fast=False):
    """nsteps device-controlled SSP33 attempts (clawhip_sharp2_ssp33_run) on the
    planar ping-pong buffers bufs = (q0, q1) (comps: their Kahan compensations
    or (None, None)); indicator: the learned indicator's path (its flags, list,
    network) with the positivity judge over the fields listed in `positive`;
    stage_events: 6*nsteps recorded torch.cuda.Events, recorded again around the
    march of stages 1-3 of every attempt (eager runs only)."""
    meqn, my, mx = bufs[0].shape
    g = L.geom2(mx, my)
    code = _dtype_code(bufs[0], fast)
    fl = lst = cnt = params = None
    flat = thr = 0.0
    mask = 0
    if indicator is not None:
        fl, (lst, cnt) = indicator._flags, indicator.flag_list
        params, flat, thr = indicator.params, indicator.flat, indicator.threshold
        for f in positive:
            mask |= 1 << int(f)
    rc = L.lib().clawhip_sharp2_ssp33_run(
        ctypes.byref(rp), code, int(char_decomp), ctypes.byref(g), L.ints(bc_lower),
        L.ints(bc_upper), L.ints(reflect), float(dx), float(dy), _ptr(bufs[0]), _ptr(bufs[1]),
        _ptr(comps[0]), _ptr(comps[1]), _ptr(dq_dt), _ptr(r2), _ptr(aux23), _ptr(fl), _ptr(lst),
        _ptr(cnt), _ptr(recon), _ptr(dq_dt_full), _ptr(params), float(flat), float(thr), mask,
        _ptr(ctl_dev), int(nsteps), _event_handles(stage_events, 6 * int(nsteps)), _stream())
    L.check(rc, "clawhip_sharp2_ssp33_run")


def _event_handles(events, n):
    """A C array of the hipEvent_t handles of n recorded torch.cuda.Events, or None."""
    if events is None:
        return None
    if len(events) < n:
        raise ValueError("stage_events needs %d events" % n)
    arr = (ctypes.c_void_p * n)()
    for i in range(n):
        h = int(events[i].cuda_event)
        if h == 0:
            raise ValueError("stage events must have been recorded once (their handle exists)")
        arr[i] = h
    return arr


def lmm_update(mode, q, u, coeffs, out, R=None, v=None, fast=False):
    """Euler / RK / SSPLMMk2 / SSPLMMk3 / LMM updates (include/clawhip.h,
    clawhip_lmm_update; q is read by modes 0, 1, 2 and 5)."""
    c = (ctypes.c_double * 5)(*[float(x) for x in (list(coeffs) + [0.0] * 5)[:5]])
    for t in (u, R, v, out):
        assert t is None or (t.shape == q.shape and t.dtype == q.dtype and t.is_contiguous())
    rc = L.lib().clawhip_lmm_update(_dtype_code(q, fast), int(mode), int(q.numel()), _ptr(q), _ptr(u),
                                    _ptr(R), _ptr(v), c, _ptr(out), _stream())
    L.check(rc, "clawhip_lmm_update")


def ssp104_mid(q, s1, fast=False):
    rc = L.lib().clawhip_ssp104_mid(_dtype_code(q, fast), _ptr(q), _ptr(s1), int(q.numel()), _stream())
    L.check(rc, "clawhip_ssp104_mid")


def step2_unsplit_run(rp, bufs, comps, ctl_dev, bc_lower, bc_upper, reflect, method, mthlim,
                      dx, dy, nsteps, num_ghost=2, fast=False):
    """nsteps device-controlled attempts (clawhip_step2_unsplit_run)."""
    meqn, my, mx = bufs[0].shape
    g = L.geom2(mx, my)
    code = _dtype_code(bufs[0], fast)
    rc = L.lib().clawhip_step2_unsplit_run(
        ctypes.byref(rp), code, ctypes.byref(g), L.ints(bc_lower), L.ints(bc_upper),
        L.ints(reflect), int(num_ghost), L.ints(method),
        L.ints(list(mthlim) + [0] * (4 - len(mthlim))), float(dx), float(dy), _ptr(bufs[0]),
        _ptr(bufs[1]), _ptr(comps[0]), _ptr(comps[1]), _ptr(ctl_dev), int(nsteps), _stream())
    L.check(rc, "clawhip_step2_unsplit_run")


# ---- learned troubled-interface indicator
def mlp_param_count():
    return int(L.lib().clawhip_mlp_param_count())


def mlp_num_interfaces(mx, my):
    return (mx + 1) * my + mx * (my + 1)


def mlp_features(q, bc_lower, bc_upper, reflect, flat, wtol, feat, label):
    """Features (n, 8) float32 and labels (n,) uint8 of every (interface, component)
    of the planar state q (meqn, my, mx)."""
    meqn, my, mx = q.shape
    n = mlp_num_interfaces(mx, my) * meqn
    if feat.numel() < n * 8 or label.numel() < n or feat.dtype != torch.float32 \
            or label.dtype != torch.uint8:
        raise ValueError("feat must hold (n, 8) float32 and label (n,) uint8")
    g = L.geom2(mx, my, 0)
    rc = L.lib().clawhip_mlp_features(_dtype_code(q), _ptr(q), ctypes.byref(g), L.ints(bc_lower),
                                      L.ints(bc_upper), L.ints(reflect), meqn, float(flat),
                                      float(wtol), _ptr(feat), _ptr(label), _stream())
    L.check(rc, "clawhip_mlp_features")
    return n


def mlp_train_step(feat, label, batch_idx, params, grad, adam_m, adam_v, step, lr, pos_weight,
                   loss):
    if batch_idx.dtype != torch.int32 or label.numel() * 8 > feat.numel():
        raise ValueError("batch_idx must be int32 and feat hold 8 floats per label")
    rc = L.lib().clawhip_mlp_train_step(_ptr(feat), int(label.numel()), _ptr(label),
                                        _ptr(batch_idx),
                                        int(batch_idx.numel()), _ptr(params), _ptr(grad),
                                        _ptr(adam_m), _ptr(adam_v), int(step), float(lr),
                                        float(pos_weight), _ptr(loss), _stream())
    L.check(rc, "clawhip_mlp_train_step")


def mlp_logits(feat, n, params, logits):
    rc = L.lib().clawhip_mlp_logits(_ptr(feat), int(n), _ptr(params), _ptr(logits), _stream())
    L.check(rc, "clawhip_mlp_logits")


def recon_elems(mx, my, meqn):
    """Elements of the work buffer a flagged stage stores its full-WENO5 edge
    values in: 2*meqn per interface."""
    return 2 * meqn * mlp_num_interfaces(mx, my)


def mlp_flag_list(flags, mx, my, out=None):
    """(ids int32, count int32 (1,)) of the interfaces a flagged stage
    reconstructs with the full WENO5 (clawhip_mlp_flag_list)."""
    n = mlp_num_interfaces(mx, my)
    if out is None or out[0].numel() < n:
        out = (torch.empty(n, dtype=torch.int32, device=flags.device),
               torch.zeros(1, dtype=torch.int32, device=flags.device))
    rc = L.lib().clawhip_mlp_flag_list(_ptr(flags), int(mx), int(my), _ptr(out[0]),
                                       _ptr(out[1]), _stream())
    L.check(rc, "clawhip_mlp_flag_list")
    return out


def mlp_flags(q, bc_lower, bc_upper, reflect, flat, params, threshold, flags):
    meqn, my, mx = q.shape
    if flags.numel() < mlp_num_interfaces(mx, my) or flags.dtype != torch.uint8:
        raise ValueError("flags must hold (mx+1)*my + mx*(my+1) uint8")
    g = L.geom2(mx, my, 0)
    rc = L.lib().clawhip_mlp_flags(_dtype_code(q), _ptr(q), ctypes.byref(g), L.ints(bc_lower),
                                   L.ints(bc_upper), L.ints(reflect), meqn, float(flat),
                                   _ptr(params), float(threshold), _ptr(flags), _stream())
    L.check(rc, "clawhip_mlp_flags")


# ---- FP64 reductions over a state
_sums_work = {}


def _work(nfields, device):
    n = int(L.lib().clawhip_sums_work_doubles(int(nfields)))
    key = (str(device), n)
    w = _sums_work.get(key)
    if w is None:
        w = torch.empty(n, dtype=torch.float64, device=device)
        _sums_work[key] = w
    return w


def state_sums(q, positive=(), out=None):
    """Device float64 tensor (meqn + 1,): the FP64 total of each field of the
    planar state q ((meqn, my, mx) or (meqn, mz, my, mx)) over its cells, then the
    number of cells where a field listed in `positive` is <= 0 or not finite
    (clawhip_state_sums; fixed reduction order)."""
    meqn = q.shape[0]
    g = L.geom2(q.shape[2], q.shape[1]) if q.dim() == 3 else \
        L.geom3(q.shape[3], q.shape[2], q.shape[1])
    if out is None:
        out = torch.empty(meqn + 1, dtype=torch.float64, device=q.device)
    mask = 0
    for f in positive:
        mask |= 1 << int(f)
    rc = L.lib().clawhip_state_sums(_dtype_code(q), _ptr(q), ctypes.byref(g), meqn, mask,
                                    _ptr(out), _ptr(_work(meqn + 1, q.device)), _stream())
    L.check(rc, "clawhip_state_sums")
    return out


def l1_diff(a, b, out=None):
    """Device float64 tensor (2,): sum |a - b| and sum |b| (clawhip_l1_diff)."""
    if a.shape != b.shape or a.dtype != b.dtype:
        raise ValueError("l1_diff needs two arrays of one shape and dtype")
    if out is None:
        out = torch.empty(2, dtype=torch.float64, device=a.device)
    rc = L.lib().clawhip_l1_diff(_dtype_code(a), _ptr(a), _ptr(b), int(a.numel()), _ptr(out),
                                 _ptr(_work(2, a.device)), _stream())
    L.check(rc, "clawhip_l1_diff")
    return out
