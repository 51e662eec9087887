"""Frame files: the binary reader (pyclaw/fileio/binary.py:21-208) on the
reference's own regression frame (src/pyclaw/tests/test_data/advection_2d_binary,
an AMRClaw frame of 17 patches on 3 levels with 2 ghost cells, used by
src/pyclaw/tests/test_io.py:24-27), and the read -> write ascii -> read round
trip of test_io.py:37-63 with util.check_solutions_are_same's checks
(util.py:286-316).

The binary frame is the advection_2d initial condition (q = 0.9 on
0.1 < x, y < 0.6, else 0.1), so every patch's interior must equal that
function at its own cell centres: a wrong ghost width, axis order, patch
offset or byte order fails it.
"""
import os

import numpy as np
import pytest

SRC = "advection_2d_binary"


def _qinit(patch):
    x, y = [patch.lower_global[k] + (np.arange(patch.num_cells_global[k]) + 0.5) *
            patch.delta[k] for k in range(2)]
    X, Y = np.meshgrid(x, y, indexing="ij")
    return 0.9 * (0.1 < X) * (X < 0.6) * (0.1 < Y) * (Y < 0.6) + 0.1


def _read(path, fmt, frame=0, **kw):
    import pyclaw_amd as pyclaw
    sol = pyclaw.Solution()
    sol.read(frame, path=path, file_format=fmt, **kw)
    return sol


def _same(a, b, qtol):
    """util.check_solutions_are_same (util.py:286-316)."""
    assert len(a.states) == len(b.states)
    assert a.t == b.t
    for s in a.states:
        r = [x for x in b.states if x.patch.patch_index == s.patch.patch_index][0]
        assert np.linalg.norm(s.q - r.q) < qtol
        if r.aux is not None:
            assert np.linalg.norm(s.aux - r.aux) < 1e-16
        for attr in ("t", "num_eqn", "num_aux"):
            assert getattr(s, attr) == getattr(r, attr)
        assert s.patch.level == r.patch.level
        for d, rd in zip(s.patch.dimensions, r.patch.dimensions):
            assert (d.num_cells, d.lower, d.delta) == (rd.num_cells, rd.lower, rd.delta)


def test_binary_reads_reference_amr_frame(golden_dir):
    src = os.path.join(golden_dir, SRC)
    sol = _read(src, "binary")
    assert sol.t == 0.0 and len(sol.states) == 17
    levels = sorted(s.patch.level for s in sol.states)
    assert levels.count(1) == 1 and set(levels) == {1, 2, 3}
    assert sorted(s.patch.patch_index for s in sol.states) == list(range(1, 18))
    cells = 0
    for s in sol.states:
        assert s.q.shape == (1,) + tuple(s.patch.num_cells_global)
        np.testing.assert_array_equal(s.q[0], _qinit(s.patch))
        cells += int(np.prod([n + 4 for n in s.patch.num_cells_global]))
    # the data file holds exactly the patches with their 2 ghost cells per side
    assert os.path.getsize(os.path.join(src, "fort.b0000")) == 8 * cells
    coarse = [s for s in sol.states if s.patch.level == 1][0]
    assert coarse.patch.num_cells_global == [50, 50] and coarse.patch.delta == [0.02, 0.02]


def test_binary_to_ascii_round_trip(golden_dir, tmp_path):
    """test_io.py:37-63 with the ascii format (HDF5 needs h5py, absent here)."""
    import pyclaw_amd as pyclaw
    ref = _read(os.path.join(golden_dir, SRC), "binary")
    ref.write(0, path=str(tmp_path), file_format="ascii")
    back = _read(str(tmp_path), "ascii")
    _same(back, ref, 1e-6)
    # the ascii header names the format, so file_format=None reads it too
    again = pyclaw.Solution(0, path=str(tmp_path))
    _same(again, ref, 1e-6)


def test_binary32_and_aux(golden_dir, tmp_path):
    """binary32 frames and aux: the reference frame rewritten as float32, and a
    synthetic 2-field aux file with ghosts whose interior is a known function;
    frame 3 falls back to fort.a0000 as binary.py:156-166 does."""
    src = os.path.join(golden_dir, SRC)
    with open(os.path.join(src, "fort.t0000")) as f:
        lines = f.readlines()
    lines[3] = lines[3].replace("    0  ", "    2  ", 1)          # naux 0 -> 2
    assert lines[3].split()[0] == "2"
    for frame in (0, 3):
        with open(tmp_path / ("fort.t%04d" % frame), "w") as f:
            f.writelines(lines)
        with open(os.path.join(src, "fort.q0000")) as f, \
                open(tmp_path / ("fort.q%04d" % frame), "w") as g:
            g.write(f.read())
        np.fromfile(os.path.join(src, "fort.b0000"), dtype=np.float64).astype(
            np.float32).tofile(tmp_path / ("fort.b%04d" % frame))
    ref = _read(src, "binary")
    blocks, expect = [], []
    for s in ref.states:
        mx, my = s.patch.num_cells_global
        gx = np.arange(-2, mx + 2)[:, None] + 1000.0 * np.arange(-2, my + 2)[None, :]
        full = np.stack([gx + s.patch.patch_index, -gx * s.patch.level])
        blocks.append(full.ravel(order="F"))
        expect.append(full[:, 2:-2, 2:-2])
    np.concatenate(blocks).astype(np.float32).tofile(tmp_path / "fort.a0000")
    for frame in (0, 3):
        sol = _read(str(tmp_path), "binary32", frame=frame, read_aux=True)
        assert len(sol.states) == 17
        for s, r, e in zip(sol.states, ref.states, expect):
            np.testing.assert_array_equal(s.q, r.q.astype(np.float32))
            assert s.aux.shape == (2,) + tuple(s.patch.num_cells_global)
            np.testing.assert_array_equal(s.aux, e.astype(np.float32))
    nan_aux = _read(str(tmp_path), "binary32", read_aux=False)
    assert np.isnan(nan_aux.states[0].aux).all()


def test_binary_short_file_and_unknown_format(golden_dir, tmp_path):
    src = os.path.join(golden_dir, SRC)
    for name in ("fort.t0000", "fort.q0000"):
        with open(os.path.join(src, name)) as f, open(tmp_path / name, "w") as g:
            g.write(f.read())
    np.fromfile(os.path.join(src, "fort.b0000"))[:1000].tofile(tmp_path / "fort.b0000")
    with pytest.raises(IOError):
        _read(str(tmp_path), "binary")
    with pytest.raises(NotImplementedError):
        _read(str(tmp_path), "hdf5")
