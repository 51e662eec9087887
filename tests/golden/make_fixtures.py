"""Convert the reference's end-to-end regression fixtures into plain ``.npz`` files.

The reference stores its expected solutions as ``expected_sols.npy``: a 0-d object
array whose payload is a pickled ``dict`` of flat float64 arrays.  No unpickler is
run on those files here.  ``pickletools.genops`` only *disassembles* the opcode
stream (it constructs nothing and imports nothing), and this script reads the
float64 payload bytes straight out of the ``BINBYTES`` opcode arguments, pairing
each payload with the dict key (``BINUNICODE``) that precedes it and checking the
element count against the shape integers that precede the payload.

Plain-text fixtures (``verify_classic.txt``) are read with ``numpy.loadtxt``.

Run in the development container only (it reads ``/root/reference``); the
outputs are committed under ``tests/golden/``.

Sources (reference test that consumes each fixture):
  examples/advection_2d/test_2d_advection.py:20-38
  examples/euler_2d/test_euler2d.py:28-37
  examples/shallow_2d/test_shallow2d.py:25-36
  examples/euler_3d/test_euler3d.py:26-57
  examples/acoustics_2d_homogeneous/test_2d_acoustics.py:19-30 (verify_classic.txt,
    verify_sharpclaw.txt, verify_sharpclaw_lmm.txt)
  examples/advection_2d_annulus/test_2d_advection_annulus.py:10-32
  examples/acoustics_2d_variable/test_acoustics_2d_variable.py:6-23 (pressure_classic.txt)
  examples/acoustics_2d_variable/test_acoustics_2d_variable_io.py:1-60 (io_test_verification)
  examples/shallow_2d/test_shallow2d.py:38-39 (sill)
  examples/psystem_2d/test_2d_psystem.py:9-25 (verify_gauge*.txt)
"""
import os
import pickletools
import shutil
import sys

import numpy as np
from numpy.lib import format as npformat

REF = "/root/reference/examples"
HERE = os.path.dirname(os.path.abspath(__file__))

WANT = {
    "advection_2d": ["classic_dim_split", "classic_unsplit_no_trans",
                     "classic_unsplit_trans_inc", "classic_unsplit_trans_cor",
                     "sharpclaw"],
    "euler_2d": ["quadrants_roe", "shock_bubble_interaction"],
    "shallow_2d": ["radialdambreak_classic_roe", "radialdambreak_sharpclaw_roe",
                   "radialdambreak_classic_hlle", "radialdambreak_sharpclaw_hlle", "sill"],
    "euler_3d": ["shocktube", "shock_bubble"],
    "advection_2d_annulus": ["classic"],
}

_DTYPE_CODES = {"O8", "f8", "|", "<", ">", "=", "i8", "f4", "i4"}


def extract_float_arrays(path):
    """Return {key: float64 array} from an object-array .npy without unpickling."""
    with open(path, "rb") as f:
        version = npformat.read_magic(f)
        if version == (1, 0):
            shape, fortran, dtype = npformat.read_array_header_1_0(f)
        else:
            shape, fortran, dtype = npformat.read_array_header_2_0(f)
        assert dtype == np.dtype("O") and shape == ()
        stream = f.read()
    out = {}
    key = None
    ints = []
    data_codes = set()
    for op, arg, _pos in pickletools.genops(stream):
        name = op.name
        if name in ("BINUNICODE", "SHORT_BINUNICODE", "UNICODE"):
            if arg in _DTYPE_CODES:
                if arg not in ("O8", "|", "<", ">", "="):
                    data_codes.add(arg)
                continue
            key, ints = arg, []
        elif name in ("BININT", "BININT1", "BININT2") and key is not None:
            ints.append(arg)
        elif name in ("BINBYTES", "SHORT_BINBYTES", "BINBYTES8") and key is not None:
            if len(arg) < 8:
                continue
            # dtype objects after the first are memo references (BINGET), so the
            # check is on the set of element codes the whole stream declares.
            assert data_codes == {"f8"}, f"{path}: element dtypes {data_codes}"
            n = len(arg) // 8
            assert n in ints, f"{path}:{key}: shape {ints} does not match {n} values"
            out[key] = np.frombuffer(arg, dtype="<f8").copy()
            key = None
    return out


def main():
    for example, keys in WANT.items():
        arrays = extract_float_arrays(os.path.join(REF, example, "expected_sols.npy"))
        missing = [k for k in keys if k not in arrays]
        if missing:
            sys.exit(f"{example}: missing keys {missing}; found {sorted(arrays)}")
        dst = os.path.join(HERE, f"{example}_expected.npz")
        np.savez_compressed(dst, **{k: arrays[k] for k in keys})
        print(dst, {k: arrays[k].shape for k in keys})
    p = np.loadtxt(os.path.join(REF, "acoustics_2d_homogeneous", "verify_classic.txt"))
    dst = os.path.join(HERE, "acoustics_2d_verify_classic.npz")
    np.savez_compressed(dst, pressure=p)
    print(dst, p.shape)
    sharp = {k: np.loadtxt(os.path.join(REF, "acoustics_2d_homogeneous", f"verify_{k}.txt"))
             for k in ("sharpclaw", "sharpclaw_lmm")}
    dst = os.path.join(HERE, "acoustics_2d_verify_sharpclaw.npz")
    np.savez_compressed(dst, **sharp)
    print(dst, {k: v.shape for k, v in sharp.items()})
    p = np.loadtxt(os.path.join(REF, "acoustics_2d_variable", "pressure_classic.txt"))
    dst = os.path.join(HERE, "acoustics_2d_variable_pressure_classic.npz")
    np.savez_compressed(dst, pressure=p)
    print(dst, p.shape)
    p = np.loadtxt(os.path.join(REF, "acoustics_2d_variable", "pressure_sharpclaw.txt"))
    dst = os.path.join(HERE, "acoustics_2d_variable_pressure_sharpclaw.npz")
    np.savez_compressed(dst, pressure=p)
    print(dst, p.shape)
    # the io test's verification frames, kept verbatim (they are ascii data)
    io_src = os.path.join(REF, "acoustics_2d_variable", "io_test_verification")
    io_dst = os.path.join(HERE, "acoustics_2d_variable_io")
    os.makedirs(io_dst, exist_ok=True)
    for name in ("fort.t0000", "fort.q0000", "fort.a0000", "fort.t0020", "fort.q0020"):
        shutil.copyfile(os.path.join(io_src, name), os.path.join(io_dst, name))
    print(io_dst)
    rho = np.loadtxt(os.path.join(REF, "advection_reaction_2d", "advection_reaction.txt"))
    dst = os.path.join(HERE, "advection_reaction_2d_expected.npz")
    np.savez_compressed(dst, density=rho)
    print(dst, rho.shape)
    # the psystem_2d example's gauge records (t, stress, 10*stress per line)
    gauges = {}
    for name in ("0.25_0.25", "17.85_1.25", "3.25_18.75", "11.75_11.75"):
        gauges["gauge" + name] = np.loadtxt(os.path.join(REF, "psystem_2d",
                                                         f"verify_gauge{name}.txt"))
    dst = os.path.join(HERE, "psystem_2d_gauges.npz")
    np.savez_compressed(dst, **gauges)
    print(dst, {k: v.shape for k, v in gauges.items()})


if __name__ == "__main__":
    main()
