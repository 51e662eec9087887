"""clawhip_state_sums_noalloc (the allocation-free FP64 reduction entry point)
against clawhip_state_sums on the same inputs, through the C ABI: two calls in a
row, the second with a smaller state, each bitwise equal to the entry point that
takes a work buffer."""
import ctypes

import numpy as np
import pytest

from pyclaw_amd import _lib as L

pytestmark = pytest.mark.gpu


def _state(rng, meqn, my, mx, dtype):
    import torch
    q = rng.uniform(0.1, 2.0, (meqn, my, mx))
    q[0, 3, 5] = -0.5           # a non-positive cell in field 0
    q[0, my - 1, mx - 1] = 0.0  # and a zero one
    return torch.from_numpy(q).to("cuda", dtype)


def _sums_both(q, positive_mask):
    import torch
    lib = L.lib()
    meqn = q.shape[0]
    code = L.F64 if q.dtype == torch.float64 else L.F32K
    g = L.geom2(q.shape[2], q.shape[1])
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    work = torch.empty(int(lib.clawhip_sums_work_doubles(meqn + 1)), dtype=torch.float64,
                       device="cuda")
    ref = torch.full((meqn + 1,), float("nan"), dtype=torch.float64, device="cuda")
    out = torch.full((meqn + 1,), float("nan"), dtype=torch.float64, device="cuda")
    rc = lib.clawhip_state_sums(code, q.data_ptr(), ctypes.byref(g), meqn, positive_mask,
                                ref.data_ptr(), work.data_ptr(), stream)
    L.check(rc, "clawhip_state_sums")
    rc = lib.clawhip_state_sums_noalloc(code, q.data_ptr(), ctypes.byref(g), meqn,
                                        positive_mask, out.data_ptr(), stream)
    L.check(rc, "clawhip_state_sums_noalloc")
    torch.cuda.synchronize()
    return ref.cpu().numpy(), out.cpu().numpy()


def test_noalloc_sums_equal_work_buffer_sums_on_two_calls():
    import torch
    rng = np.random.default_rng(77)
    # first call: FP64, 7 fields, 700 x 900 (more rows than reduction blocks)
    q1 = _state(rng, 7, 700, 900, torch.float64)
    print("noalloc call 1: f64 meqn 7 700x900", flush=True)
    ref1, out1 = _sums_both(q1, 0b1)
    assert np.array_equal(ref1, out1), (ref1, out1)
    assert ref1[-1] == 2.0
    # second call, smaller: FP32, 3 fields, 37 x 53 (fewer rows than blocks; the
    # first call's larger partials are still in the shared array)
    q2 = _state(rng, 3, 37, 53, torch.float32)
    print("noalloc call 2: f32 meqn 3 37x53", flush=True)
    ref2, out2 = _sums_both(q2, 0b101)
    assert np.array_equal(ref2, out2), (ref2, out2)
    host = q2.double().cpu().numpy()
    assert np.allclose(out2[:3], host.reshape(3, -1).sum(axis=1), rtol=1e-12, atol=1e-9)
    bad = ((host[0] <= 0) | (host[2] <= 0)).sum()
    assert out2[3] == float(bad)


def test_noalloc_sums_reject_too_many_fields():
    import torch
    lib = L.lib()
    q = torch.zeros((8, 4, 4), dtype=torch.float64, device="cuda")
    out = torch.zeros(9, dtype=torch.float64, device="cuda")
    g = L.geom2(4, 4)
    rc = lib.clawhip_state_sums_noalloc(L.F64, q.data_ptr(), ctypes.byref(g), 8, 0,
                                        out.data_ptr(), None)
    assert rc == -1
