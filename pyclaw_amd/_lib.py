"""ctypes binding of libclawhip.so (declared in include/clawhip.h).

The library is the product's compute path: there is no CPU fallback.  Loading
fails loudly when the shared object has not been built.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libclawhip.so")

F64, F32K, F64_FAST = 0, 1, 2
RP_ACOUSTICS, RP_EULER, RP_SHALLOW, RP_ADVECTION, RP_EULER3 = 1, 2, 3, 4, 5
BC_EXTRAP, BC_PERIODIC, BC_WALL, BC_GIVEN = 1, 2, 3, 4

_ERRORS = {-1: "invalid argument (size, stride or pointer)",
           -2: "unknown Riemann solver or unsupported combination",
           -3: "unsupported method flags"}


class ClawhipError(RuntimeError):
    pass


class RP(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("efix", ctypes.c_int),
                ("rho", ctypes.c_double), ("bulk", ctypes.c_double),
                ("cc", ctypes.c_double), ("zz", ctypes.c_double),
                ("gamma", ctypes.c_double), ("grav", ctypes.c_double),
                ("u", ctypes.c_double), ("v", ctypes.c_double), ("meqn", ctypes.c_int),
                ("fwave", ctypes.c_int), ("dry_tolerance", ctypes.c_double),
                ("sea_level", ctypes.c_double)]


class StepCtl(ctypes.Structure):
    """clawhip_stepctl (include/clawhip.h)."""
    _fields_ = [("t", ctypes.c_double), ("dt", ctypes.c_double), ("tstart", ctypes.c_double),
                ("tend", ctypes.c_double), ("cfl", ctypes.c_double),
                ("cfl_desired", ctypes.c_double), ("cfl_max", ctypes.c_double),
                ("dt_max", ctypes.c_double), ("cflmax", ctypes.c_double),
                ("dtmin", ctypes.c_double), ("dtmax", ctypes.c_double),
                ("dt_variable", ctypes.c_int), ("done", ctypes.c_int), ("cur", ctypes.c_int),
                ("attempts", ctypes.c_int), ("max_steps", ctypes.c_int),
                ("accepted", ctypes.c_int), ("nrejected", ctypes.c_int), ("pad0", ctypes.c_int),
                ("pad2", ctypes.c_uint), ("pad1", ctypes.c_uint),
                ("cfl_bits", ctypes.c_ulonglong)]


class SharpCtl(ctypes.Structure):
    """clawhip_sharpctl (include/clawhip.h)."""
    _fields_ = [(n, ctypes.c_double) for n in
                ("t", "dt", "dt_old", "tstart", "tend", "cfl", "cfl_desired", "cfl_max",
                 "dt_max", "cflmax", "dtmin", "dtmax")] + \
               [(n, ctypes.c_int) for n in
                ("have_dt_old", "accept_prev", "done", "cur", "attempts", "max_steps",
                 "accepted", "nrejected", "judge", "age", "every", "judge_steps", "fallbacks",
                 "run_flags", "run1", "run23", "run_fb", "pad0")] + \
               [("pad1", ctypes.c_uint), ("nbad", ctypes.c_uint),
                ("cfl1_bits", ctypes.c_ulonglong), ("cfl23_bits", ctypes.c_ulonglong),
                ("cflfb_bits", ctypes.c_ulonglong)]


class EnsMember(ctypes.Structure):
    """clawhip_ens_member"""
    _fields_ = [("mx", ctypes.c_int), ("my", ctypes.c_int),
                ("q0", ctypes.c_void_p), ("q1", ctypes.c_void_p),
                ("comp0", ctypes.c_void_p), ("comp1", ctypes.c_void_p),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double),
                ("ctl", ctypes.c_void_p)]


class Geom(ctypes.Structure):
    _fields_ = [("mx", ctypes.c_int), ("my", ctypes.c_int), ("mz", ctypes.c_int),
                ("ghosts", ctypes.c_int), ("field_stride", ctypes.c_longlong),
                ("row_stride", ctypes.c_longlong), ("slab_stride", ctypes.c_longlong)]


def geom2(mx, my, ghosts=0):
    """Planar 2D array (field, j, i) with `ghosts` layers stored on every side."""
    nx, ny = mx + 2 * ghosts, my + 2 * ghosts
    return Geom(mx, my, 1, ghosts, nx * ny, nx, nx * ny)


def geom3(mx, my, mz, ghosts=0):
    """Planar 3D array (field, k, j, i) with `ghosts` layers stored on every side."""
    nx, ny, nz = mx + 2 * ghosts, my + 2 * ghosts, mz + 2 * ghosts
    return Geom(mx, my, mz, ghosts, nx * ny * nz, nx, nx * ny)


_I = ctypes.c_int
_P = ctypes.c_void_p
_D = ctypes.c_double
_F = ctypes.c_float
_IP = ctypes.POINTER(ctypes.c_int)

_SIGS = {
    "clawhip_device_count": [_IP],
    "clawhip_cfl_reset": [_P, _P],
    "clawhip_step2_unsplit": [ctypes.POINTER(RP), _I, _P, ctypes.POINTER(Geom), _P,
                              ctypes.POINTER(Geom), _P, _P, _IP, _IP, _IP, _I, _IP, _IP,
                              _D, _D, _D, _P, ctypes.POINTER(Geom), _IP, _IP, _P, _P],
    "clawhip_step2_split": [ctypes.POINTER(RP), _I, _P, ctypes.POINTER(Geom), _P,
                            ctypes.POINTER(Geom), _P, _P, _P, _P, _IP, _IP, _IP, _I, _IP, _IP,
                            _D, _D, _D, _P, ctypes.POINTER(Geom), _IP, _IP, _P, _P],
    "clawhip_step3_unsplit_work_elems": [_I, _I, _I, _I],
    "clawhip_step3_unsplit": [ctypes.POINTER(RP), _I, _P, ctypes.POINTER(Geom), _P,
                              ctypes.POINTER(Geom), _P, _IP, _IP, _IP, _I, _IP, _IP, _D, _D, _D,
                              _D, _P, _P],
    "clawhip_step3_split": [ctypes.POINTER(RP), _I, _P, ctypes.POINTER(Geom), _P,
                            ctypes.POINTER(Geom), _P, _P, _P, _P, _IP, _IP, _IP, _I, _IP, _IP,
                            _D, _D, _D, _D, _P, _P],
    "clawhip_sharp2_stage": [ctypes.POINTER(RP), _I, _I, ctypes.POINTER(Geom), _IP, _IP, _IP,
                             _I, _I, _D, _D, _D, _D, _P, ctypes.POINTER(Geom), _P, _P, _P, _P,
                             _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, ctypes.POINTER(Geom), _IP,
                             _IP, _I, _P, _P],
    "clawhip_sharp2_ssp33_run": [ctypes.POINTER(RP), _I, _I, ctypes.POINTER(Geom), _IP, _IP, _IP,
                                 _D, _D, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                 _F, _F, ctypes.c_uint, _P, _I, _P, _P],
    "clawhip_ssp104_mid": [_I, _P, _P, ctypes.c_longlong, _P],
    "clawhip_lmm_update": [_I, _I, ctypes.c_longlong, _P, _P, _P, _P,
                           ctypes.POINTER(ctypes.c_double), _P, _P],
    "clawhip_step2_unsplit_run": [ctypes.POINTER(RP), _I, ctypes.POINTER(Geom), _IP, _IP, _IP,
                                  _I, _IP, _IP, _D, _D, _P, _P, _P, _P, _P, _I, _P],
    "clawhip_step2_sweep": [ctypes.POINTER(RP), _I, _I, _P, ctypes.POINTER(Geom), _P,
                            ctypes.POINTER(Geom), _I, _IP, _IP, _D, _D, _D, _P,
                            ctypes.POINTER(Geom), _P, _P],
    "clawhip_step3_sweep": [ctypes.POINTER(RP), _I, _I, _P, ctypes.POINTER(Geom), _P,
                            ctypes.POINTER(Geom), _I, _IP, _IP, _D, _D, _D, _D, _P, _P],
    "clawhip_sharp2_flux": [ctypes.POINTER(RP), _I, _I, _P, ctypes.POINTER(Geom), _P,
                            ctypes.POINTER(Geom), _D, _D, _D, _P, ctypes.POINTER(Geom), _P, _P],
    "clawhip_ensemble_table_bytes": [_I],
    "clawhip_ensemble_prepare": [ctypes.POINTER(RP), _I, ctypes.POINTER(EnsMember), _I, _IP,
                                 _IP, _IP, _I, _IP, _IP, _I, _P,
                                 ctypes.POINTER(ctypes.c_longlong)],
    "clawhip_ensemble_run": [ctypes.POINTER(RP), _I, _P, ctypes.POINTER(ctypes.c_longlong), _I,
                             _P],
    "clawhip_mlp_param_count": [],
    "clawhip_mlp_features": [_I, _P, ctypes.POINTER(Geom), _IP, _IP, _IP, _I, _F, _F, _P, _P,
                             _P],
    "clawhip_mlp_train_step": [_P, ctypes.c_longlong, _P, _P, _I, _P, _P, _P, _P, _I, _F, _F,
                               _P, _P],
    "clawhip_mlp_logits": [_P, ctypes.c_longlong, _P, _P, _P],
    "clawhip_mlp_flag_list": [_P, _I, _I, _P, _P, _P],
    "clawhip_mlp_flags": [_I, _P, ctypes.POINTER(Geom), _IP, _IP, _IP, _I, _F, _P, _F, _P, _P],
    "clawhip_sums_work_doubles": [_I],
    "clawhip_state_sums": [_I, _P, ctypes.POINTER(Geom), _I, ctypes.c_uint, _P, _P, _P],
    "clawhip_state_sums_noalloc": [_I, _P, ctypes.POINTER(Geom), _I, ctypes.c_uint, _P, _P],
    "clawhip_l1_diff": [_I, _P, _P, ctypes.c_longlong, _P, _P, _P],
}

_LONG_RESULT = {"clawhip_ensemble_table_bytes", "clawhip_sums_work_doubles",
                "clawhip_step3_unsplit_work_elems"}

_lib = None


def lib():
    """Load libclawhip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ClawhipError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                f"g.build()'` (hipcc, gfx950).  pyclaw_amd has no CPU fallback.")
        L = ctypes.CDLL(LIB_PATH)
        for name, argtypes in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = argtypes
            f.restype = ctypes.c_longlong if name in _LONG_RESULT else ctypes.c_int
        _lib = L
    return _lib


def check(rc, what):
    if rc != 0:
        msg = _ERRORS.get(rc, f"HIP error {rc}")
        raise ClawhipError(f"{what}: {msg}")


def ints(seq):
    return (ctypes.c_int * len(seq))(*[int(x) for x in seq])


def declared_symbols():
    """Names declared in include/clawhip.h (parsed), for the export check."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "clawhip.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"\bint\s+(clawhip_\w+)\s*\(", text)))
