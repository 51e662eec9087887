"""The classic Clawpack ascii frame files, as pyclaw/fileio/ascii.py writes and
reads them: fort.tNNNN (time and counts), fort.qNNNN (per patch a header, then
the cells with x fastest, one line of num_eqn values per cell, a blank line
after each row of a 2D patch), fort.aNNNN (aux, same layout).

Values are written "%18.8e" (ascii.py:97-126) and headers as ascii.py:84-95.
The reference also pickles problem_data and mapc2p into fort.pklNNNN; those
files are neither written nor read here (reading one would run its pickle).
"""
import os

import numpy as np

_FMT = "%18.8e"


def _frame_name(path, prefix, kind, frame):
    return os.path.join(path, "%s.%s%s" % (prefix, kind, str(frame).zfill(4)))


def write(solution, frame, path, file_prefix="fort", write_aux=False, options={},
          write_p=False):
    """ascii.py:17-81 (without the pickle side file)."""
    with open(_frame_name(path, file_prefix, "t", frame), "w") as f:
        f.write("%18.8e     time\n" % solution.t)
        f.write("%5i                  num_eqn\n" % solution.num_eqn)
        f.write("%5i                  nstates\n" % len(solution.states))
        f.write("%5i                  num_aux\n" % solution.num_aux)
        f.write("%5i                  num_dim\n" % solution.domain.num_dim)
        f.write("%5i                  num_ghost\n" % 0)
        f.write("%s                  file_format\n" % "ascii")
    with open(_frame_name(path, file_prefix, "q", frame), "w") as f:
        for state in solution.states:
            write_patch_header(f, state.patch)
            write_array(f, state.patch, state.p if write_p else state.q)
    if solution.num_aux > 0 and write_aux:
        with open(_frame_name(path, file_prefix, "a", frame), "w") as f:
            for state in solution.states:
                write_patch_header(f, state.patch)
                write_array(f, state.patch, state.aux)


def write_patch_header(f, patch):
    """ascii.py:84-95"""
    f.write("%5i                  patch_number\n" % patch.patch_index)
    f.write("%5i                  AMR_level\n" % patch.level)
    for dim in patch.dimensions:
        f.write("%5i                  m%s\n" % (dim.num_cells, dim.name))
    for dim in patch.dimensions:
        f.write("%18.8e     %slow\n" % (dim.lower, dim.name))
    for dim in patch.dimensions:
        f.write("%18.8e     d%s\n" % (dim.delta, dim.name))
    f.write("\n")


def write_array(f, patch, q):
    """ascii.py:97-126: cells with x fastest, num_eqn values per line; in 2D a
    blank line closes each row, in 3D also each slab."""
    q = np.asarray(q)
    nd = patch.num_dim
    if nd == 1:
        np.savetxt(f, q.T, fmt=_FMT, delimiter="")
    elif nd == 2:
        for j in range(q.shape[2]):
            np.savetxt(f, q[:, :, j].T, fmt=_FMT, delimiter="")
            f.write("\n")
    elif nd == 3:
        for k in range(q.shape[3]):
            for j in range(q.shape[2]):
                np.savetxt(f, q[:, :, j, k].T, fmt=_FMT, delimiter="")
                f.write("\n")
            f.write("\n")
    else:
        raise Exception("Dimension Exception in writing fort file.")


def _data_line(f, kind=float):
    """util.read_data_line: the first token of the next line."""
    line = f.readline()
    tokens = line.split()
    if not tokens:
        raise ValueError("empty header line")
    return kind(tokens[0])


def read_t(frame, path="./", file_prefix="fort"):
    """ascii.py:242-297: (t, num_eqn, nstates, num_aux, num_dim, num_ghost,
    file_format); the last two default to 0 and None in older files."""
    with open(_frame_name(path, file_prefix, "t", frame), "r") as f:
        t = _data_line(f)
        num_eqn = _data_line(f, int)
        nstates = _data_line(f, int)
        num_aux = _data_line(f, int)
        num_dim = _data_line(f, int)
        try:
            num_ghost = _data_line(f, int)
        except ValueError:
            num_ghost = 0
        try:
            file_format = _data_line(f, str)
        except ValueError:
            file_format = None
    return t, num_eqn, nstates, num_aux, num_dim, num_ghost, file_format


def read_patch_header(f, num_dim):
    """ascii.py:300-339: a Patch with dimensions named x, y, z."""
    from ..geometry import Dimension, Patch
    patch_index = _data_line(f, int)
    level = _data_line(f, int)
    n = [_data_line(f, int) for _ in range(num_dim)]
    lower = [_data_line(f) for _ in range(num_dim)]
    d = [_data_line(f) for _ in range(num_dim)]
    f.readline()
    names = ["x", "y", "z"]
    patch = Patch([Dimension(lower[i], lower[i] + n[i] * d[i], n[i], name=names[i])
                   for i in range(num_dim)])
    patch.patch_index = patch_index
    patch.level = level
    return patch


def read_array(f, patch, num_var):
    """ascii.py:342-400: the values of each cell may span several lines; blank
    lines close rows (2D) and slabs (3D)."""
    shape = [num_var] + list(patch.num_cells_global)
    q = np.zeros(shape, order="F")
    cells = int(np.prod(shape[1:]))
    vals = []
    for _ in range(cells):
        row = []
        while len(row) < num_var:
            line = f.readline()
            if line == "":
                raise IOError("Unexpected EOF in %s" % f.name)
            row += line.split()
        vals.append([float(x) for x in row[:num_var]])
        # a blank line follows each completed row / slab
        done = len(vals)
        if patch.num_dim >= 2 and done % shape[1] == 0:
            f.readline()
            if patch.num_dim == 3 and done % (shape[1] * shape[2]) == 0:
                f.readline()
    arr = np.array(vals).T          # (num_var, cells), x fastest
    q[...] = arr.reshape([num_var] + shape[1:][::-1]).transpose(
        [0] + list(range(len(shape) - 1, 0, -1)))
    return q


def read(solution, frame, path="./", file_prefix="fort", read_aux=False, options={}):
    """ascii.py:129-239 (no pickle side file): the frame's states, patches and
    domain, aux from fort.aNNNN or else fort.a0000."""
    from ..geometry import Domain
    from ..state import State
    t, num_eqn, nstates, num_aux, num_dim, _, _ = read_t(frame, path, file_prefix)
    patches = []
    with open(_frame_name(path, file_prefix, "q", frame), "r") as f:
        for _ in range(nstates):
            patch = read_patch_header(f, num_dim)
            state = State(patch, num_eqn, num_aux)
            state.t = t
            if num_aux > 0:
                state.aux[:] = np.nan
            state.q = read_array(f, patch, num_eqn)
            solution.states.append(state)
            patches.append(patch)
    solution.domain = Domain(patches)
    if num_aux > 0 and read_aux:
        fname = _frame_name(path, file_prefix, "a", frame)
        if not os.path.exists(fname):
            fname = _frame_name(path, file_prefix, "a", 0)
            if not os.path.exists(fname):
                return
        with open(fname, "r") as f:
            for state in solution.states:
                aux_patch = read_patch_header(f, num_dim)
                for a, b in zip(state.patch.dimensions, aux_patch.dimensions):
                    assert a.num_cells == b.num_cells, "aux file header does not match"
                    assert abs(a.lower - b.lower) <= 1e-8 + 1e-15 * abs(a.lower)
                    assert abs(a.delta - b.delta) <= 1e-8 + 1e-15 * abs(a.delta)
                state.aux = read_array(f, state.patch, num_aux)
