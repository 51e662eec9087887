"""f2py-compatible stand-in for clawpack.pyclaw.sharpclaw.sharpclaw2 (flux2 and
the clawparams / workspace / reconstruct modules SharpClawSolver sets up,
sharpclaw/solver.py:563-588) running the HIP kernels."""
import ctypes

import numpy as np

from .. import _lib as L
from . import CParam, new_cfl, ptr, rp_struct, stream, to_device2

cparam = CParam()


class _ClawParams(object):
    """Attributes of the Fortran module ClawParams (sharpclaw/ClawParams.f90)."""

    def __init__(self):
        self.num_dim = 2
        self.lim_type = 2
        self.weno_order = 5
        self.char_decomp = 0
        self.tfluct_solver = False
        self.fwave = False
        self.index_capa = 0
        self.num_waves = 1
        self.xlower = None
        self.xupper = None
        self.dx = None
        self.mthlim = None

    def alloc_clawparams(self):
        self.xlower = np.zeros(self.num_dim)
        self.xupper = np.zeros(self.num_dim)

    def dealloc_clawparams(self):
        pass


class _Workspace(object):
    def alloc_workspace(self, maxnx, num_ghost, num_eqn, num_waves, char_decomp):
        pass

    def dealloc_workspace(self, char_decomp):
        pass


class _Reconstruct(object):
    def alloc_recon_workspace(self, maxnx, num_ghost, num_eqn, num_waves, lim_type, char_decomp):
        pass

    def dealloc_recon_workspace(self, lim_type, char_decomp):
        pass


clawparams = _ClawParams()
workspace = _Workspace()
reconstruct = _Reconstruct()


def flux2(q, aux, dt, t, num_ghost, maxm, mx, my, rpn2, tfluct2):
    """dq = dt*L(q) (sharpclaw/flux2.f90:2-74) on the interior of a qbc whose
    ghost cells the caller filled.  Returns (dq, cfl), dq shaped like q."""
    p = clawparams
    if p.lim_type != 2 or p.weno_order != 5 or p.tfluct_solver or p.fwave:
        raise NotImplementedError("pyclaw_amd implements lim_type=2, weno_order=5, no tfluct/fwave")
    if p.char_decomp not in (0, 2) or int(p.index_capa) != 0:
        raise NotImplementedError("char_decomp 0 or 2, no capacity function")
    if num_ghost < 3:
        raise ValueError("WENO5 needs 3 ghost cells")
    rp = rp_struct(rpn2, cparam, meqn=q.shape[0])
    dx, dy = [float(d) for d in p.dx]
    gin = L.geom2(mx, my, num_ghost)
    gout = L.geom2(mx, my, 0)
    qin = to_device2(q)
    out = qin.new_empty((q.shape[0], my, mx))
    cfl = new_cfl()
    auxd = to_device2(aux) if (aux is not None and np.size(aux) > 1 and
                               getattr(rp, "kind", 0) in (6, 7, 8)) else None
    rc = L.lib().clawhip_sharp2_flux(ctypes.byref(rp), L.F64, int(p.char_decomp), ptr(qin),
                                     ctypes.byref(gin), ptr(out), ctypes.byref(gout), float(dt),
                                     dx, dy, ptr(auxd) if auxd is not None else None,
                                     ctypes.byref(gin) if auxd is not None else None,
                                     ptr(cfl), stream())
    L.check(rc, "clawhip_sharp2_flux")
    dq = np.zeros_like(q, order="F")
    dq[:, num_ghost:num_ghost + mx, num_ghost:num_ghost + my] = np.transpose(out.cpu().numpy(),
                                                                             (0, 2, 1))
    return dq, float(cfl.item())
