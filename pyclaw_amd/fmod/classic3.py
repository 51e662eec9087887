"""f2py-compatible stand-in for clawpack.pyclaw.classic.classic3 (step3ds,
step3) running the HIP kernels.  Set ``solver.fmod`` to this module."""
import ctypes

import torch

from .. import _lib as L
from . import (CParam, check_method, from_device3, mthlim4, new_cfl, ptr, rp_struct, stream,
               to_device3)

cparam = CParam()


def step3ds(maxm, num_ghost, mx, my, mz, qold, qnew, aux, dx, dy, dz, dt, method, mthlim, aux1,
            aux2, aux3, work, idir, use_fwave, rpn3, rpt3, rptt3):
    """One dimension-split sweep (classic/step3ds.f90:2-378): idir 1 = x, 2 = y,
    3 = z.  Returns (qnew, cfl)."""
    check_method(method, use_fwave)
    rp = rp_struct(rpn3, cparam)
    g = L.geom3(mx, my, mz, num_ghost)
    qin = to_device3(qold)
    qout = torch.empty_like(qin) if idir == 1 else qin
    cfl = new_cfl()
    rc = L.lib().clawhip_step3_sweep(ctypes.byref(rp), L.F64, int(idir), ptr(qin),
                                     ctypes.byref(g), ptr(qout), ctypes.byref(g), int(num_ghost),
                                     L.ints(method), L.ints(mthlim4(mthlim)), float(dx), float(dy),
                                     float(dz), float(dt), ptr(cfl), stream())
    L.check(rc, "clawhip_step3_sweep")
    qnew[...] = from_device3(qout)
    return qnew, float(cfl.item())


def step3(maxm, num_ghost, mx, my, mz, qold, qnew, aux, dx, dy, dz, dt, method, mthlim, aux1,
          aux2, aux3, work, use_fwave, rpn3, rpt3, rptt3):
    """The unsplit step (classic/step3.f90:2-597) on a caller-filled qbc: qnew's
    interior <- the step of qold.  qnew's ghost cells keep their values (the
    reference's step3 also writes the first ghost layer; ClawSolver3D reads only
    the interior back).  Returns (qnew, cfl)."""
    from .. import kernels as K
    check_method(method, use_fwave)
    rp = rp_struct(rpn3, cparam)
    qin = to_device3(qold)
    out = qin.new_empty((qold.shape[0], mz, my, mx))
    cfl = new_cfl()
    K.step3_unsplit(rp, qin, out, cfl, [4, 4, 4], [4, 4, 4], [1, 2, 3], method,
                    mthlim4(mthlim), dx, dy, dz, dt, num_ghost=num_ghost,
                    gin=L.geom3(mx, my, mz, num_ghost))
    ng = num_ghost
    qnew[:, ng:ng + mx, ng:ng + my, ng:ng + mz] = from_device3(out)
    return qnew, float(cfl.item())
