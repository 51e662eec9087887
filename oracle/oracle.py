"""ctypes bindings for the oracle (oracle/liboracle.so) and, where it has been
built, the reference's own compiled Fortran step (oracle/_ref/*.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()`` and
by ``bench.py``'s cpu_baseline leg, as the checker -- never by the product
package ``pyclaw_amd``.

All arrays use the reference layout: numpy float64, Fortran order, equation
index first, ghost cells included, e.g. ``qbc.shape == (meqn, mx+2ng, my+2ng)``.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_CLASSIC_PATH = os.path.join(HERE, "_ref", "libref_classic.so")
REF_SHARP_PATH = os.path.join(HERE, "_ref", "libref_sharp.so")

ACOUSTICS, EULER, SHALLOW, ADVECTION, EULER3 = 1, 2, 3, 4, 5
CUSTOM, EXTRAP, PERIODIC, WALL = 0, 1, 2, 3

MEQN = {ACOUSTICS: 3, EULER: 4, SHALLOW: 3, ADVECTION: 1, EULER3: 5, 8: 3, 9: 3, 10: 5, 11: 3}
MWAVES = {ACOUSTICS: 2, EULER: 4, SHALLOW: 3, ADVECTION: 1, EULER3: 3, 6: 1, 7: 1, 8: 2, 9: 2, 10: 5,
          11: 3}
VC_ADVECTION = 6
VC_ADVECTION_2D = 7
VC_ACOUSTICS = 8
SHALLOW_HLLE = 9
EULER5 = 10
SHALLOW_FWAVE = 11


class RP(ctypes.Structure):
    """Mirror of ``orc_rp`` (oracle.h)."""
    _fields_ = [("kind", ctypes.c_int), ("efix", ctypes.c_int),
                ("rho", ctypes.c_double), ("bulk", ctypes.c_double),
                ("cc", ctypes.c_double), ("zz", ctypes.c_double),
                ("gamma", ctypes.c_double), ("grav", ctypes.c_double),
                ("u", ctypes.c_double), ("v", ctypes.c_double), ("meqn", ctypes.c_int),
                ("dry_tolerance", ctypes.c_double), ("sea_level", ctypes.c_double)]


def make_rp(kind, efix=1, rho=1.0, bulk=4.0, gamma=1.4, grav=1.0, u=0.5, v=1.0, meqn=0,
            dry_tolerance=1e-3, sea_level=0.0):
    cc = float(np.sqrt(bulk / rho))
    return RP(kind, efix, rho, bulk, cc, rho * cc, gamma, grav, u, v, meqn, dry_tolerance,
              sea_level)


_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def _dp(a):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1)
    return a.ctypes.data_as(_D)


def _ip(seq):
    arr = (ctypes.c_int * len(seq))(*[int(x) for x in seq])
    return arr


_lib = None
_ref_classic = None
_ref_sharp = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -f oracle/Makefile)")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_step2.restype = ctypes.c_double
        L.orc_step2ds.restype = ctypes.c_double
        L.orc_step3ds.restype = ctypes.c_double
        L.orc_sharp_flux2.restype = ctypes.c_double
        L.orc_sharp_flux1.restype = ctypes.c_double
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_CLASSIC_PATH) and os.path.exists(REF_SHARP_PATH)


def ref_classic():
    global _ref_classic
    if _ref_classic is None:
        L = ctypes.CDLL(REF_CLASSIC_PATH)
        L.ref_step2.restype = ctypes.c_double
        L.ref_step3ds.restype = ctypes.c_double
        _ref_classic = L
    return _ref_classic


def ref_sharp():
    global _ref_sharp
    if _ref_sharp is None:
        L = ctypes.CDLL(REF_SHARP_PATH)
        L.ref_sharp_flux2.restype = ctypes.c_double
        L.ref_sharp_flux1.restype = ctypes.c_double
        _ref_sharp = L
    return _ref_sharp


def method_array(order=2, transverse=2, dimensional_split=False, dt_variable=True, num_aux=0,
                 index_capa=-1):
    """ClawSolver._set_method (classic/solver.py:173-193) for 2D/3D; index_capa is
    the 0-based capacity field (-1: none), stored 1-based in method(6)."""
    m = [int(dt_variable), order, -1 if dimensional_split else transverse, 0, 0,
         int(index_capa) + 1, int(num_aux)]
    return m


# ---------------------------------------------------------------- BCs
def bc2(qbc, ng, bc_lower, bc_upper, reflect=(1, 2)):
    meqn, nx, ny = qbc.shape
    lib().orc_bc2(_dp(qbc), meqn, nx - 2 * ng, ny - 2 * ng, ng, _ip(bc_lower), _ip(bc_upper),
                  _ip(reflect))
    return qbc


def bc3(qbc, ng, bc_lower, bc_upper, reflect=(1, 2, 3)):
    meqn, nx, ny, nz = qbc.shape
    lib().orc_bc3(_dp(qbc), meqn, nx - 2 * ng, ny - 2 * ng, nz - 2 * ng, ng, _ip(bc_lower),
                  _ip(bc_upper), _ip(reflect))
    return qbc


def with_ghosts(q, ng):
    """Interior q (meqn, mx, my[, mz]) -> zero-padded qbc in Fortran order."""
    shape = (q.shape[0],) + tuple(n + 2 * ng for n in q.shape[1:])
    qbc = np.zeros(shape, order="F")
    sl = (slice(None),) + tuple(slice(ng, -ng) for _ in q.shape[1:])
    qbc[sl] = q
    return qbc


def interior(qbc, ng):
    sl = (slice(None),) + tuple(slice(ng, -ng) for _ in qbc.shape[1:])
    return np.asfortranarray(qbc[sl])


# ---------------------------------------------------------------- classic
def step2(rp, qold, qnew, ng, dx, dy, dt, method, mthlim, ids=0, use_fwave=False, aux=None):
    """ids=0: unsplit step2; ids=1/2: step2ds sweep.  qnew updated in place.
    aux: aux(maux, 1-ng:mx+ng, 1-ng:my+ng) with ghosts filled; method[5] =
    index_capa (1-based) names a capacity field (unsplit)."""
    meqn, nx, ny = qold.shape
    mx, my = nx - 2 * ng, ny - 2 * ng
    mw = MWAVES[rp.kind]
    maux = 0 if aux is None else aux.shape[0]
    if ids == 0:
        return lib().orc_step2(meqn, mw, ng, mx, my, _dp(qold), _dp(qnew), ctypes.c_double(dx),
                               ctypes.c_double(dy), ctypes.c_double(dt), _ip(method),
                               _ip(mthlim), int(use_fwave), ctypes.byref(rp), int(maux),
                               _dp(aux) if aux is not None else None)
    return lib().orc_step2ds(meqn, mw, ng, mx, my, _dp(qold), _dp(qnew), ctypes.c_double(dx),
                             ctypes.c_double(dy), ctypes.c_double(dt), _ip(method), _ip(mthlim),
                             int(ids), int(use_fwave), ctypes.byref(rp), int(maux),
                             _dp(aux) if aux is not None else None)


def step3(rp, qold, qnew, ng, dx, dy, dz, dt, method, mthlim, use_fwave=False):
    """The unsplit 3D step (classic/step3.f90); qnew updated in place."""
    meqn, nx, ny, nz = qold.shape
    f = lib().orc_step3
    f.restype = ctypes.c_double
    return f(meqn, MWAVES[rp.kind], ng, nx - 2 * ng, ny - 2 * ng, nz - 2 * ng, _dp(qold),
             _dp(qnew), ctypes.c_double(dx), ctypes.c_double(dy), ctypes.c_double(dz),
             ctypes.c_double(dt), _ip(method), _ip(mthlim), int(use_fwave), ctypes.byref(rp))


def rpt3(rp, ixyz, icoor, qL, qR, asdq):
    """The oracle's euler_3D transverse split of asdq at the interface qL | qR."""
    bm, bp = np.zeros(5), np.zeros(5)
    lib().orc_rpt3(ctypes.byref(rp), int(ixyz), int(icoor), _dp(np.ascontiguousarray(qL, float)),
                   _dp(np.ascontiguousarray(qR, float)), _dp(np.ascontiguousarray(asdq, float)),
                   _dp(bm), _dp(bp))
    return bm, bp


def step3ds(rp, qold, qnew, ng, dx, dy, dz, dt, method, mthlim, idir, use_fwave=False):
    meqn, nx, ny, nz = qold.shape
    return lib().orc_step3ds(meqn, MWAVES[rp.kind], ng, nx - 2 * ng, ny - 2 * ng, nz - 2 * ng,
                             _dp(qold), _dp(qnew), ctypes.c_double(dx), ctypes.c_double(dy),
                             ctypes.c_double(dz), ctypes.c_double(dt), _ip(method), _ip(mthlim),
                             int(idir), int(use_fwave), ctypes.byref(rp))


# ---------------------------------------------------------------- SharpClaw
def sharp_flux2(rp, q, ng, dx, dy, dt, char_decomp, aux=None):
    """aux: aux(maux, 1-ng:mx+ng, 1-ng:my+ng) with ghosts filled, or None."""
    meqn, nx, ny = q.shape
    dq = np.zeros_like(q, order="F")
    maux = 0 if aux is None else aux.shape[0]
    cfl = lib().orc_sharp_flux2(meqn, MWAVES[rp.kind], ng, nx - 2 * ng, ny - 2 * ng, _dp(q),
                                _dp(dq), ctypes.c_double(dx), ctypes.c_double(dy),
                                ctypes.c_double(dt), int(char_decomp), ctypes.byref(rp),
                                int(maux), _dp(aux) if aux is not None else None)
    return dq, cfl


def set_weno_eps(eps):
    """The WENO epsilon of the restatement's weno5 / weno5_char (test hook; the
    default is the reference's REAL(4) 1.e-36, (double)1.0e-36f)."""
    lib().orc_set_weno_eps(ctypes.c_double(eps))


REF_WENO_EPS = float(np.float32(1.0e-36))


def sharp_flux1(rp, q1d, ng, ixy, dtdx, char_decomp):
    meqn, n2 = q1d.shape
    dq1d = np.zeros_like(q1d, order="F")
    cfl = lib().orc_sharp_flux1(meqn, MWAVES[rp.kind], ng, n2 - 2 * ng, ixy, _dp(q1d),
                                _dp(dq1d), ctypes.c_double(dtdx), int(char_decomp),
                                ctypes.byref(rp), 0, None)
    return dq1d, cfl


def weno5(q1d, ng):
    """Component-wise WENO5 edge values (orc_weno5, weno.f90:5-83) of a
    1D slice (meqn, n2) with ng ghosts: returns (ql, qr), left and right edges."""
    q1d = np.asfortranarray(q1d, dtype=np.float64)
    meqn, n2 = q1d.shape
    ql = np.zeros_like(q1d, order="F")
    qr = np.zeros_like(q1d, order="F")
    lib().orc_weno5(meqn, ng, n2, _dp(q1d), _dp(ql), _dp(qr))
    return ql, qr


def evec(rp, ixy, q1d):
    meqn, n2 = q1d.shape
    evl = np.zeros((meqn, meqn, n2), order="F")
    evr = np.zeros((meqn, meqn, n2), order="F")
    lib().orc_evec(ctypes.byref(rp), ixy, meqn, n2, _dp(q1d), _dp(evl), _dp(evr))
    return evl, evr


def rpn(rp, ixy, qL, qR):
    meqn, mw = MEQN[rp.kind], MWAVES[rp.kind]
    wave = np.zeros(meqn * mw)
    s = np.zeros(mw)
    amdq = np.zeros(meqn)
    apdq = np.zeros(meqn)
    qL = np.ascontiguousarray(qL, dtype=np.float64)
    qR = np.ascontiguousarray(qR, dtype=np.float64)
    lib().orc_rpn(ctypes.byref(rp), ixy, _dp(qL), _dp(qR), _dp(wave), _dp(s), _dp(amdq),
                  _dp(apdq))
    return wave.reshape(mw, meqn), s, amdq, apdq


def rpn_aux(rp, ixy, qL, qR, auxL, auxR):
    """The normal solve of the solvers that read aux (orc_rpn_aux)."""
    meqn, mw = MEQN[rp.kind], MWAVES[rp.kind]
    wave = np.zeros(meqn * mw)
    s = np.zeros(mw)
    amdq = np.zeros(meqn)
    apdq = np.zeros(meqn)
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (qL, qR, auxL, auxR)]
    lib().orc_rpn_aux(ctypes.byref(rp), ixy, *[_dp(a) for a in args], _dp(wave), _dp(s),
                      _dp(amdq), _dp(apdq))
    return wave.reshape(mw, meqn), s, amdq, apdq


# ---------------------------------------------------------------- reference build
def ref_fwave(on):
    """use_fwave for the reference step routines called through the shim."""
    ref_classic().shim_set_fwave(int(bool(on)))


def ref_step3(rp, qold, qnew, ng, dx, dy, dz, dt, method, mthlim):
    """The reference's own unsplit step3 with the oracle's euler_3D solvers."""
    meqn, nx, ny, nz = qold.shape
    f = ref_classic().ref_step3
    f.restype = ctypes.c_double
    return f(ctypes.byref(rp), meqn, MWAVES[rp.kind], ng, nx - 2 * ng, ny - 2 * ng, nz - 2 * ng,
             _dp(qold), _dp(qnew), ctypes.c_double(dx), ctypes.c_double(dy), ctypes.c_double(dz),
             ctypes.c_double(dt), _ip(method), _ip(mthlim))


def ref_step2(rp, qold, qnew, ng, dx, dy, dt, method, mthlim, ids=0, use_fwave=False):
    meqn, nx, ny = qold.shape
    ref_fwave(use_fwave)
    try:
        return ref_classic().ref_step2(ctypes.byref(rp), int(ids), meqn, MWAVES[rp.kind], ng,
                                       nx - 2 * ng, ny - 2 * ng, _dp(qold), _dp(qnew),
                                       ctypes.c_double(dx), ctypes.c_double(dy),
                                       ctypes.c_double(dt), _ip(method), _ip(mthlim))
    finally:
        ref_fwave(False)


_ref_vcadv = None
REF_VCADV_PATH = os.path.join(os.path.dirname(REF_CLASSIC_PATH), "libref_vcadv.so")


def ref_vcadv_rpn2():
    """Address of the advection_reaction_2d example's rpn2 (built from its source)."""
    global _ref_vcadv
    if _ref_vcadv is None:
        _ref_vcadv = ctypes.CDLL(REF_VCADV_PATH)
    return ctypes.cast(_ref_vcadv.rpn2_, ctypes.c_void_p)


def ref_step2ds_aux(rpn2, qold, qnew, aux, ng, dx, dy, dt, method, mthlim, ids, mwaves=1,
                    rp=None, use_fwave=False):
    """The reference's step2ds sweep with aux fields and the given rpn2 address
    (None: the oracle's solver rp)."""
    meqn, nx, ny = qold.shape
    if rp is not None:
        ref_classic().shim_set_rp(ctypes.byref(rp))
    f = ref_classic().ref_step2ds_aux
    f.restype = ctypes.c_double
    ref_fwave(use_fwave)
    try:
        return f(rpn2, int(ids), meqn, int(mwaves), aux.shape[0], ng, nx - 2 * ng,
                 ny - 2 * ng, _dp(qold), _dp(qnew), _dp(aux), ctypes.c_double(dx),
                 ctypes.c_double(dy), ctypes.c_double(dt), _ip(method), _ip(mthlim))
    finally:
        ref_fwave(False)


_ref_euler5 = None
REF_EULER5_PATH = os.path.join(os.path.dirname(REF_CLASSIC_PATH), "libref_euler5.so")


def ref_euler5(gamma):
    """Addresses of the reference tree's own euler_5wave_2D rpn2/rpt2 (built from
    development/rp_approaches, behind an adapter to step2's argument lists),
    with their /cparam/ gamma set."""
    global _ref_euler5
    if _ref_euler5 is None:
        _ref_euler5 = ctypes.CDLL(REF_EULER5_PATH)
        _ref_euler5.euler5_set_gamma.argtypes = [ctypes.c_double]
    _ref_euler5.euler5_set_gamma(float(gamma))
    return (ctypes.cast(_ref_euler5.euler5_rpn2_, ctypes.c_void_p),
            ctypes.cast(_ref_euler5.euler5_rpt2_, ctypes.c_void_p))


def ref_step2_fns(rp, rpn2, rpt2, qold, qnew, ng, dx, dy, dt, method, mthlim, ids=0):
    """The reference's step2 (ids 0) or step2ds sweep (1, 2) with the given
    Riemann solver addresses (None: the oracle's for rp); no aux fields."""
    meqn, nx, ny = qold.shape
    f = ref_classic().ref_step2_fns
    f.restype = ctypes.c_double
    return f(ctypes.byref(rp), rpn2, rpt2, int(ids), meqn, MWAVES[rp.kind], ng, nx - 2 * ng,
             ny - 2 * ng, _dp(qold), _dp(qnew), ctypes.c_double(dx), ctypes.c_double(dy),
             ctypes.c_double(dt), _ip(method), _ip(mthlim))


def ref_step2_aux(rp, rpn2, qold, qnew, aux, ng, dx, dy, dt, method, mthlim, mwaves=1):
    """The reference's unsplit step2 with aux fields (and a capacity field when
    method[5] > 0), the given rpn2 address and the oracle's rpt2 for rp."""
    meqn, nx, ny = qold.shape
    f = ref_classic().ref_step2_aux
    f.restype = ctypes.c_double
    return f(ctypes.byref(rp), rpn2, meqn, int(mwaves), aux.shape[0], ng, nx - 2 * ng,
             ny - 2 * ng, _dp(qold), _dp(qnew), _dp(aux), ctypes.c_double(dx),
             ctypes.c_double(dy), ctypes.c_double(dt), _ip(method), _ip(mthlim))


def ref_step3ds(rp, qold, qnew, ng, dx, dy, dz, dt, method, mthlim, idir):
    meqn, nx, ny, nz = qold.shape
    return ref_classic().ref_step3ds(ctypes.byref(rp), int(idir), meqn, MWAVES[rp.kind], ng,
                                     nx - 2 * ng, ny - 2 * ng, nz - 2 * ng, _dp(qold),
                                     _dp(qnew), ctypes.c_double(dx), ctypes.c_double(dy),
                                     ctypes.c_double(dz), ctypes.c_double(dt), _ip(method),
                                     _ip(mthlim))


def ref_sharp_flux2(rp, q, ng, dx, dy, dt, char_decomp, aux=None, rpn2=None):
    """The reference's SharpClaw flux2 (aux: ghost-filled aux fields or None; the
    Riemann solver with aux is the oracle's).  rpn2: the address of a solver with
    the reference's argument list to use instead of the oracle's (no aux)."""
    meqn, nx, ny = q.shape
    dq = np.zeros_like(q, order="F")
    maux = 0 if aux is None else aux.shape[0]
    L = ref_sharp()
    L.shim_set_sharp_rpn2(rpn2)
    try:
        cfl = L.ref_sharp_flux2(ctypes.byref(rp), int(char_decomp), meqn,
                                MWAVES[rp.kind], ng, nx - 2 * ng, ny - 2 * ng, _dp(q),
                                _dp(dq), ctypes.c_double(dx), ctypes.c_double(dy),
                                ctypes.c_double(dt), int(maux),
                                _dp(aux) if aux is not None else None)
    finally:
        L.shim_set_sharp_rpn2(None)
    return dq, cfl


def ref_sharp_flux1(rp, q1d, ng, ixy, dxd, dt, char_decomp):
    meqn, n2 = q1d.shape
    dq1d = np.zeros_like(q1d, order="F")
    cfl = ref_sharp().ref_sharp_flux1(ctypes.byref(rp), int(char_decomp), meqn,
                                      MWAVES[rp.kind], ng, n2 - 2 * ng, int(ixy), _dp(q1d),
                                      _dp(dq1d), ctypes.c_double(dxd), ctypes.c_double(dt))
    return dq1d, cfl
