"""Device storage of states: torch tensors (HIP memory) in planar layout.

A state q of the reference's shape (meqn, mx, my[, mz]) is stored as a
contiguous tensor of shape (meqn, [mz,] my, mx): field-major planes with the x
index fastest, so a wavefront's 64 lanes read 64 consecutive cells of a row.
"""
import numpy as np
import torch


def torch_dtype(dtype):
    return torch.float64 if dtype == 0 else torch.float32


def to_planar(q):
    """(meqn, mx, my[, mz]) -> contiguous (meqn, [mz,] my, mx) numpy array."""
    axes = (0,) + tuple(range(q.ndim - 1, 0, -1))
    return np.ascontiguousarray(np.transpose(q, axes))


def from_planar(a):
    """(meqn, [mz,] my, mx) -> Fortran-ordered (meqn, mx, my[, mz]) numpy array."""
    axes = (0,) + tuple(range(a.ndim - 1, 0, -1))
    return np.asfortranarray(np.transpose(a, axes))


def q_to_device(q, dtype=0, device="cuda"):
    return torch.from_numpy(to_planar(np.asarray(q, dtype=np.float64))).to(
        device=device, dtype=torch_dtype(dtype))


def q_from_device(t):
    return from_planar(t.detach().to("cpu", torch.float64).numpy())


def stream_handle():
    return ctypes_void_p(torch.cuda.current_stream().cuda_stream)


def ctypes_void_p(x):
    import ctypes
    return ctypes.c_void_p(int(x))


class no_gc(object):
    """Context: collect garbage now, then keep the collector off until exit.  Used
    around graph capture, where the collector could otherwise free an earlier
    graph (its destructor calls HIP, which a capturing stream refuses)."""

    def __enter__(self):
        import gc
        gc.collect()
        self._was = gc.isenabled()
        gc.disable()
        return self

    def __exit__(self, *exc):
        import gc
        if self._was:
            gc.enable()
        return False
