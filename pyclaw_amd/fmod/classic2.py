"""f2py-compatible stand-in for clawpack.pyclaw.classic.classic2 (step2,
step2ds) running the HIP kernels.  Set ``solver.fmod`` to this module."""
import ctypes

import torch

from .. import _lib as L
from .. import kernels as K
from . import (CParam, check_method, from_device2, mthlim4, new_cfl, ptr, rp_struct, stream,
               to_device2)

cparam = CParam()


def step2(maxm, num_ghost, mx, my, qold, qnew, aux, dx, dy, dt, method, mthlim, aux1, aux2,
          aux3, work, use_fwave, rpn2, rpt2):
    """Unsplit step (classic/step2.f90:2-250): qnew's interior <- step of qold
    (ghost cells filled by the caller, in aux too).  Returns (qnew, cfl)."""
    check_method(method, use_fwave, allow_aux=True, allow_capa=True, allow_fwave=True)
    rp = rp_struct(rpn2, cparam, meqn=qold.shape[0], use_fwave=use_fwave)
    qin = to_device2(qold)
    out = qin.new_empty((qold.shape[0], my, mx))
    cfl = new_cfl()
    auxd = to_device2(aux) if (aux is not None and method[6] > 0) else None
    K.step2_unsplit(rp, qin, out, cfl, [4, 4], [4, 4], [1, 2], method, mthlim4(mthlim), dx, dy,
                    dt, num_ghost=num_ghost, gin=L.geom2(mx, my, num_ghost), aux=auxd,
                    aux_bc_lower=[4, 4], aux_bc_upper=[4, 4])
    qnew[:, num_ghost:num_ghost + mx, num_ghost:num_ghost + my] = from_device2(out)
    return qnew, float(cfl.item())


def step2ds(maxm, num_ghost, mx, my, qold, qnew, aux, dx, dy, dt, method, mthlim, aux1, aux2,
            aux3, work, ids, use_fwave, rpn2, rpt2):
    """One dimension-split sweep (classic/step2ds.f90:2-264): ids 1 = x, 2 = y.
    Returns (qnew, cfl)."""
    check_method(method, use_fwave, allow_aux=True, allow_fwave=True)
    rp = rp_struct(rpn2, cparam, meqn=qold.shape[0], use_fwave=use_fwave)
    g = L.geom2(mx, my, num_ghost)
    qin = to_device2(qold)
    qout = torch.empty_like(qin) if ids == 1 else qin
    cfl = new_cfl()
    auxd = to_device2(aux) if (aux is not None and method[6] > 0) else None
    rc = L.lib().clawhip_step2_sweep(ctypes.byref(rp), L.F64, int(ids), ptr(qin), ctypes.byref(g),
                                     ptr(qout), ctypes.byref(g), int(num_ghost), L.ints(method),
                                     L.ints(mthlim4(mthlim)), float(dx), float(dy), float(dt),
                                     ptr(auxd) if auxd is not None else None,
                                     ctypes.byref(g) if auxd is not None else None,
                                     ptr(cfl), stream())
    L.check(rc, "clawhip_step2_sweep")
    qnew[...] = from_device2(qout)
    return qnew, float(cfl.item())
