"""The C-ABI library (include/clawhip.h) loads without a GPU and exports every
declared entry point; host-side validation of the Python wrappers."""
import ctypes

import numpy as np
import pytest

from pyclaw_amd import _lib as L


def test_library_exports_match_header():
    lib = L.lib()
    names = L.declared_symbols()
    assert len(names) >= 13
    for name in names:
        assert hasattr(lib, name), name
    # every declared entry point has a ctypes signature
    assert set(names) <= set(L._SIGS), set(names) - set(L._SIGS)


def test_ctypes_signatures_match_header_arity():
    """Every ctypes signature has as many arguments as the header declares."""
    import os
    import re
    hdr = open(os.path.join(os.path.dirname(L.HERE), "include", "clawhip.h")).read()
    for name, sig in L._SIGS.items():
        m = re.search(r"\b" + name + r"\s*\(([^)]*)\)", hdr)
        assert m, name
        args = [a for a in m.group(1).split(",") if a.strip() and a.strip() != "void"]
        assert len(args) == len(sig), (name, len(args), len(sig))


def test_noalloc_sums_reject_sizes_before_device_work():
    """The allocation-free reduction refuses more fields than its fixed scratch
    holds (and null pointers) before it touches a device."""
    lib = L.lib()
    g = L.geom2(4, 4)
    dummy = ctypes.c_void_p(16)
    assert lib.clawhip_state_sums_noalloc(L.F64, dummy, ctypes.byref(g), 8, 0, dummy, None) == -1
    assert lib.clawhip_state_sums_noalloc(L.F64, dummy, ctypes.byref(g), 0, 0, dummy, None) == -1
    assert lib.clawhip_state_sums_noalloc(L.F64, None, ctypes.byref(g), 3, 0, dummy, None) == -1


def test_mlp_param_count():
    # W1[16][32] (row 6 = bias), W2[32][32], b2[32], w3[32], b3
    assert L.lib().clawhip_mlp_param_count() == 16 * 32 + 32 * 32 + 32 + 32 + 1


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "libclawhip.so"))
    with pytest.raises(L.ClawhipError):
        L.lib()


def _small_solution(pyclaw, meqn, num_aux=0):
    x = pyclaw.Dimension(0.0, 1.0, 8, name="x")
    y = pyclaw.Dimension(0.0, 1.0, 8, name="y")
    domain = pyclaw.Domain([x, y])
    state = pyclaw.State(domain, meqn, num_aux)
    state.q[...] = 1.0
    return pyclaw.Solution(state, domain)


def test_unsupported_options_raise_at_setup():
    """Options the device solvers do not implement fail loudly at setup instead of
    being ignored (no GPU needed: setup raises before any device work)."""
    import pyclaw_amd as pyclaw
    from pyclaw_amd import riemann
    s = pyclaw.SharpClawSolver2D(riemann.shallow_roe_with_efix_2D)
    s.call_before_step_each_stage = True
    s.before_step = lambda solver, state: None
    with pytest.raises(NotImplementedError):
        s.setup(_small_solution(pyclaw, 3))
    s = pyclaw.SharpClawSolver2D(riemann.shallow_roe_with_efix_2D)
    s.time_integrator = "LMM"
    s.alpha, s.beta = [0.0, 1.0], [0.0, 1.0]
    s.cfl_max = 0.5
    with pytest.raises(AssertionError):   # LMM needs a fixed dt (sharpclaw/solver.py:222-224)
        s.setup(_small_solution(pyclaw, 3))
    s = pyclaw.SharpClawSolver2D(riemann.shallow_roe_with_efix_2D)
    s.time_integrator = "RK"
    s.a, s.b = [[0.0]], [1.0]
    with pytest.raises(KeyError):         # RK has no default CFL (sharpclaw/solver.py:246-247)
        s.setup(_small_solution(pyclaw, 3))
    c = pyclaw.ClawSolver3D(riemann.euler_3D)   # f-waves: the 2D steps only
    c.fwave = True
    sol3 = pyclaw.Solution(5, pyclaw.Domain([0., 0., 0.], [1., 1., 1.], [4, 4, 4]))
    with pytest.raises(NotImplementedError):
        c.setup(sol3)
    s = pyclaw.SharpClawSolver2D(riemann.shallow_roe_with_efix_2D)
    s.fwave = True
    with pytest.raises(NotImplementedError):
        s.setup(_small_solution(pyclaw, 3))
    c = pyclaw.ClawSolver2D(riemann.shallow_bathymetry_fwave_2D)   # no transverse solver
    c.dimensional_split = False
    sol = _small_solution(pyclaw, 3)
    sol.state.aux = np.zeros((1,) + sol.state.q.shape[1:], order="F")
    with pytest.raises(NotImplementedError):
        c.setup(sol)
