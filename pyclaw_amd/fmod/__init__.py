"""Drop-in kernel modules for the reference's own solver classes.

PyClaw's ClawSolver and SharpClawSolver import their compiled kernel module only
when ``solver.fmod`` is unset (classic/solver.py:208-211, sharpclaw/solver.py:
254-257).  Setting

    solver.fmod = pyclaw_amd.fmod.classic2      # ClawSolver2D
    solver.fmod = pyclaw_amd.fmod.classic3      # ClawSolver3D
    solver.fmod = pyclaw_amd.fmod.sharpclaw2    # SharpClawSolver2D

routes step2 / step2ds / step3ds / flux2 to the HIP kernels with the f2py
signatures the solvers call (classic/solver.py:527-541, 680-690;
sharpclaw/solver.py:893).  The Riemann solver must be one of pyclaw_amd.riemann's
modules: its ``rpn2._cpointer`` is the device solver's tag.  The arrays cross
PCIe on every call (the reference's numpy qbc in and out); pyclaw_amd's own
solvers keep q resident in HBM instead.
"""
import ctypes

import numpy as np
import torch

from .. import kernels as K


class CParam(object):
    """Stands for the Riemann solver's /cparam/ common block: State.set_cparam
    (state.py:206-231) copies problem_data onto it."""


def rp_struct(rpn, cparam, meqn=0, use_fwave=False):
    """Device Riemann-solver parameters from the solver tag and cparam (meqn: the
    number of fields, for the solvers that take any; use_fwave: step2's flag)."""
    if isinstance(rpn, bool) or not isinstance(rpn, (int, np.integer)):
        raise TypeError("pyclaw_amd.fmod runs the device Riemann solvers: use a "
                        "pyclaw_amd.riemann module as the solver's riemann_solver")
    pd = {k: v for k, v in vars(cparam).items() if not k.startswith("_")}
    return K.make_rp(int(rpn), pd, meqn=meqn, fwave=use_fwave)


def check_method(method, use_fwave, allow_aux=False, allow_capa=False, allow_fwave=False):
    if use_fwave and not allow_fwave:
        raise NotImplementedError("f-wave Riemann solvers are supported by the 2D steps")
    if int(method[5]) != 0 and not allow_capa:
        raise NotImplementedError("capacity functions are supported by the unsplit 2D step only")
    if int(method[6]) != 0 and not allow_aux:
        raise NotImplementedError("aux arrays are supported by the 2D steps only")


def mthlim4(mthlim):
    m = [int(x) for x in np.atleast_1d(mthlim)]
    return m + [0] * (4 - len(m))


def to_device2(qbc):
    """F-order (meqn, nx, ny) host array -> planar (meqn, ny, nx) FP64 device tensor."""
    return torch.from_numpy(np.ascontiguousarray(np.transpose(qbc, (0, 2, 1)),
                                                 dtype=np.float64)).cuda()


def to_device3(qbc):
    return torch.from_numpy(np.ascontiguousarray(np.transpose(qbc, (0, 3, 2, 1)),
                                                 dtype=np.float64)).cuda()


def from_device2(t):
    return np.transpose(t.cpu().numpy(), (0, 2, 1))


def from_device3(t):
    return np.transpose(t.cpu().numpy(), (0, 3, 2, 1))


def new_cfl():
    return torch.zeros(1, dtype=torch.float64, device="cuda")


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
