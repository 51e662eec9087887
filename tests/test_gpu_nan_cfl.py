"""A state with a NaN cell (one that overflowed): the device steps leave NaN
wave speeds out of the Courant number exactly as the oracle does (the rule is
pinned on the CPU in tests/test_oracle.py), and the NaNs land in the same cells."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.test_oracle import nan_cell_cfls, nan_cell_state

pytestmark = pytest.mark.gpu


def test_nan_cell_courant_numbers_match_oracle():
    from pyclaw_amd import kernels as K
    from pyclaw_amd.device import q_from_device, q_to_device
    q = nan_cell_state()
    mx, my = q.shape[1:]
    dx, dy = 1.0 / mx, 1.0 / my
    dt = 0.1 * dx
    c_unsplit, c_xsweep, c_sharp0, c_sharp2 = nan_cell_cfls(q, 3)
    rp = K.make_rp(O.EULER, {"gamma": 1.4})
    orp = O.make_rp(O.EULER, efix=1, gamma=1.4)
    # unsplit step: Courant number and state
    qbc = O.with_ghosts(q, 2)
    O.bc2(qbc, 2, [1, 1], [1, 1])
    qnew = qbc.copy(order="F")
    O.step2(orp, qbc, qnew, 2, dx, dy, dt, O.method_array(transverse=2), [4] * 4)
    qi = q_to_device(q)
    qo = torch.empty_like(qi)
    cfl = torch.zeros(1, dtype=torch.float64, device="cuda")
    print("nan cfl: unsplit", flush=True)
    K.step2_unsplit(rp, qi, qo, cfl, [1, 1], [1, 1], [1, 2], O.method_array(transverse=2),
                    [4] * 4, dx, dy, dt)
    torch.cuda.synchronize()
    assert float(cfl.item()) == c_unsplit
    assert np.array_equal(q_from_device(qo), O.interior(qnew, 2), equal_nan=True)
    # dimension-split step: both sweeps
    meth = O.method_array(dimensional_split=True)
    qbc = O.with_ghosts(q, 2)
    O.bc2(qbc, 2, [1, 1], [1, 1])
    qold = qbc.copy(order="F")
    c1 = O.step2(orp, qold, qbc, 2, dx, dy, dt, meth, [4] * 4, ids=1)
    c2 = O.step2(orp, qbc, qbc, 2, dx, dy, dt, meth, [4] * 4, ids=2)
    assert c1 == c_xsweep
    cfl.zero_()
    print("nan cfl: split", flush=True)
    K.step2_split(rp, qi, qo, cfl, [1, 1], [1, 1], [1, 2], meth, [4] * 4, dx, dy, dt)
    torch.cuda.synchronize()
    assert float(cfl.item()) == max(c1, c2)
    assert np.array_equal(q_from_device(qo), O.interior(qbc, 2), equal_nan=True)
    # SharpClaw operator, component and characteristic WENO5
    for cd, cref in ((0, c_sharp0), (2, c_sharp2)):
        cfl.zero_()
        out = torch.empty_like(qi)
        print("nan cfl: sharp", cd, flush=True)
        K.sharp2_stage(rp, cd, out, cfl, [1, 1], [1, 1], [1, 2], dt, dx, dy, qi)
        torch.cuda.synchronize()
        assert float(cfl.item()) == cref, (cd, float(cfl.item()), cref)
