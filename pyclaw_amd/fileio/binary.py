"""Raw binary frames as the Clawpack Fortran codes write them (fort.tNNNN,
fort.qNNNN patch headers, fort.bNNNN data, fort.aNNNN aux), read into a
Solution -- pyclaw/fileio/binary.py:21-208.  PyClaw writes no binary frames,
so there is no writer here either.

fort.bNNNN holds every patch in fort.qNNNN's order, each as a Fortran-ordered
(num_eqn, mx + 2*num_ghost[, my + 2*num_ghost[, ...]]) block of float64
('binary', 'binary64') or float32 ('binary32') values, ghost cells included;
the reader keeps the interior.  The patch headers are the ascii format's.
"""
import os

import numpy as np

from . import ascii as A


def _dtype(options):
    fmt = options.get("format", "binary64")
    if fmt in ("binary", "binary64"):
        return np.float64
    if fmt == "binary32":
        return np.float32
    raise Exception("Unrecognized file_format: %s" % fmt)


def _interior(block, num_ghost):
    if num_ghost == 0:
        return block
    cut = (slice(None),) + (slice(num_ghost, -num_ghost),) * (block.ndim - 1)
    return block[cut]


def _blocks(data, shapes, num_ghost, what):
    """Split the flat data into per-patch Fortran-ordered blocks (interiors)."""
    out, start = [], 0
    for shape in shapes:
        full = [shape[0]] + [n + 2 * num_ghost for n in shape[1:]]
        size = int(np.prod(full))
        if start + size > data.size:
            raise IOError("%s holds %d values, its patches need more" % (what, data.size))
        out.append(_interior(data[start:start + size].reshape(full, order="F"), num_ghost))
        start += size
    return out


def read(solution, frame, path="./", file_prefix="fort", read_aux=False, options={}):
    """binary.py:21-208: the frame's patches (from fort.qNNNN's headers) and
    their q (from fort.bNNNN); with read_aux, aux from fort.aNNNN, else
    fort.a0000.  Aux values stay NaN when no aux file is read."""
    from ..geometry import Domain
    from ..state import State
    if options.get("format", "binary64") == "petsc":
        raise NotImplementedError("PETSc binary frames are not read")
    dtype = _dtype(options)
    t, num_eqn, nstates, num_aux, num_dim, num_ghost, _ = A.read_t(frame, path, file_prefix)
    if num_dim not in (1, 2, 3):
        raise Exception("Read only supported up to 3d.")
    q_name = A._frame_name(path, file_prefix, "q", frame)
    b_name = A._frame_name(path, file_prefix, "b", frame)
    qdata = np.fromfile(b_name, dtype=dtype)
    patches = []
    with open(q_name, "r") as f:
        for _ in range(nstates):
            patches.append(A.read_patch_header(f, num_dim))
    shapes = [[num_eqn] + list(p.num_cells_global) for p in patches]
    for patch, q in zip(patches, _blocks(qdata, shapes, num_ghost, b_name)):
        state = State(patch, num_eqn, num_aux)
        state.t = t
        if num_aux > 0:
            state.aux[:] = np.nan
        state.q = np.asfortranarray(q, dtype=np.float64)
        solution.states.append(state)
    solution.domain = Domain(patches)

    if not (num_aux > 0 and read_aux):
        return
    fname = A._frame_name(path, file_prefix, "a", frame)
    if not os.path.exists(fname):
        fname = A._frame_name(path, file_prefix, "a", 0)
        if not os.path.exists(fname):
            return
    auxdata = np.fromfile(fname, dtype=dtype)
    shapes = [[num_aux] + list(s.patch.num_cells_global) for s in solution.states]
    for state, aux in zip(solution.states, _blocks(auxdata, shapes, num_ghost, fname)):
        state.aux = np.asfortranarray(aux, dtype=np.float64)
