"""The C-ABI library (include/clawhip.h) loads without a GPU and exports every
declared entry point; host-side validation of the Python wrappers."""
import ctypes

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


def test_mlp_param_count():
    # W1[16][32] (row 6 = bias), W2[32][32], b2[32], w3[32], b3
    assert L.lib().clawhip_mlp_param_count() == 16 * 32 + 32 * 32 + 32 + 32 + 1


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "libclawhip.so"))
    with pytest.raises(L.ClawhipError):
        L.lib()
