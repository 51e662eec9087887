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


def _dtype_code(t):
    if t.dtype == torch.float64:
        return L.F64
    if t.dtype == torch.float32:
        return L.F32K
    raise L.ClawhipError(f"unsupported dtype {t.dtype}")


def make_rp(kind, problem_data, efix=1):
    """Parameters of the Riemann solver from problem_data (state.py:206-231)."""
    pd = problem_data
    rho = float(pd.get("rho", 1.0))
    bulk = float(pd.get("bulk", 1.0))
    cc = float(pd.get("cc", (bulk / rho) ** 0.5))
    zz = float(pd.get("zz", rho * cc))
    return L.RP(kind, int(pd.get("efix", efix)), rho, bulk, cc, zz,
                float(pd.get("gamma", 1.4)), float(pd.get("grav", 1.0)),
                float(pd.get("u", 0.0)), float(pd.get("v", 0.0)))


def cfl_reset(cfl):
    L.check(L.lib().clawhip_cfl_reset(_ptr(cfl), _stream()), "clawhip_cfl_reset")


def step2_unsplit(rp, q_in, q_out, cfl, bc_lower, bc_upper, reflect, method, mthlim,
                  dx, dy, dt, comp_in=None, comp_out=None, num_ghost=2, gin=None):
    """One unsplit step q_in -> q_out (planar (meqn, my, mx) tensors)."""
    meqn, my, mx = q_out.shape
    gout = L.geom2(mx, my)
    if gin is None:
        gin = gout
    code = _dtype_code(q_out)
    assert q_in.dtype == q_out.dtype and q_in.is_contiguous() and q_out.is_contiguous()
    rc = L.lib().clawhip_step2_unsplit(
        ctypes.byref(rp), code, _ptr(q_in), ctypes.byref(gin), _ptr(q_out), ctypes.byref(gout),
        _ptr(comp_in), _ptr(comp_out), L.ints(bc_lower), L.ints(bc_upper), L.ints(reflect), int(num_ghost),
        L.ints(method), L.ints(list(mthlim) + [0] * (4 - len(mthlim))), float(dx), float(dy),
        float(dt), _ptr(cfl), _stream())
    L.check(rc, "clawhip_step2_unsplit")


def _work_like(q_out, ng, work):
    """Padded intermediate (meqn, [mz+2ng,] my+2ng, mx+2ng), reused between calls."""
    shape = (q_out.shape[0],) + tuple(n + 2 * ng for n in q_out.shape[1:])
    if work is None or tuple(work.shape) != shape or work.dtype != q_out.dtype:
        work = torch.empty(shape, dtype=q_out.dtype, device=q_out.device)
    return work


def step2_split(rp, q_in, q_out, cfl, bc_lower, bc_upper, reflect, method, mthlim, dx, dy, dt,
                comp_in=None, comp_out=None, num_ghost=2, gin=None, work=None):
    """Both dimension-split sweeps (step2ds ids=1, 2); returns the work buffers."""
    meqn, my, mx = q_out.shape
    gout = L.geom2(mx, my)
    if gin is None:
        gin = gout
    code = _dtype_code(q_out)
    w, cw = work if work is not None else (None, None)
    w = _work_like(q_out, num_ghost, w)
    if code == L.F32K:
        cw = _work_like(q_out, num_ghost, cw)
    rc = L.lib().clawhip_step2_split(
        ctypes.byref(rp), code, _ptr(q_in), ctypes.byref(gin), _ptr(q_out), ctypes.byref(gout),
        _ptr(comp_in), _ptr(comp_out), _ptr(w), _ptr(cw), L.ints(bc_lower), L.ints(bc_upper),
        L.ints(reflect), int(num_ghost), L.ints(method),
        L.ints(list(mthlim) + [0] * (4 - len(mthlim))), float(dx), float(dy), float(dt),
        _ptr(cfl), _stream())
    L.check(rc, "clawhip_step2_split")
    return (w, cw)


def step3_split(rp, q_in, q_out, cfl, bc_lower, bc_upper, reflect, method, mthlim, dx, dy, dz,
                dt, comp_in=None, comp_out=None, num_ghost=2, work=None):
    """The three dimension-split sweeps of ClawSolver3D (step3ds idir=1,2,3)."""
    meqn, mz, my, mx = q_out.shape
    g = L.geom3(mx, my, mz)
    code = _dtype_code(q_out)
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
