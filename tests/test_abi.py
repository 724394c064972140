"""CPU checks of the C ABI: the library loads and exports every declared symbol.

No compute call is made (there is no GPU here); creation-time validation
errors are covered by the GPU tests.
"""

import ctypes
import os

import pytest

from apple_amd import _lib


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _lib.exported_symbols()
    assert "apl_elem_eval" in names and "apl_mesh_create" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_string():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = _lib.lib()
    assert L.apl_version() >= 10000
    assert isinstance(L.apl_last_error(), bytes)


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libapple_amd.so")
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.lib()


def test_setup_mode_is_validated():
    import pytest

    from apple_amd import fem
    from apple_amd.mesh import cube

    with pytest.raises(ValueError, match="setup"):
        fem.StableNeoHookean.from_pyvista(cube(2), setup="gpu")
