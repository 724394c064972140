"""CPU checks of the C ABI: the library loads and exports every declared symbol.

No compute call is made (there is no GPU here); validation that precedes
any device call (mesh size limits) is checked here, the rest of the
creation-time validation by the GPU tests.
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


@pytest.mark.parametrize("n_tets,n_verts", [(1 << 27, 8), (4, 1 << 27), (-1, 8)])
def test_mesh_size_limits_rejected_before_any_device_call(n_tets, n_verts):
    """At most 2^27 tets and vertices (32-bit byte offsets in the kernels):
    rejected with a clear error before the cells are read or a device is
    touched."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = _lib.lib()
    cells = (ctypes.c_int32 * 4)(0, 1, 2, 3)
    dummy = (ctypes.c_double * 9)()  # never read: the size check comes first
    p = lambda a: ctypes.cast(a, ctypes.c_void_p)
    out = ctypes.c_void_p()
    for create in (L.apl_mesh_create, L.apl_mesh_create_from_points):
        rc = create(_lib.APL_F32, n_tets, n_verts, p(cells), p(dummy), p(dummy), 1, None, None, 1, 0,
                    ctypes.byref(out))
        assert rc == -1 and out.value is None
        assert b"2^27" in L.apl_last_error()


def test_fp64_tet_count_limit_of_32bit_element_offsets():
    """FP64 meshes: the inverse rest shapes' 9 SoA rows are addressed by 32-bit
    byte offsets, so at most 2^26 - 1024 tets; rejected before any device call."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = _lib.lib()
    cells = (ctypes.c_int32 * 4)(0, 1, 2, 3)
    dummy = (ctypes.c_double * 9)()
    p = lambda a: ctypes.cast(a, ctypes.c_void_p)
    out = ctypes.c_void_p()
    rc = L.apl_mesh_create(_lib.APL_F64, (1 << 26) - 1023, 8, p(cells), p(dummy), p(dummy), 1, None, None, 1, 0,
                           ctypes.byref(out))
    assert rc == -1 and out.value is None
    assert b"2^26" in L.apl_last_error()


def test_layout_mode_switch_is_validated():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = _lib.lib()
    assert L.apl_layout_mode(2) == -1
    assert "layout mode" in L.apl_last_error().decode()
    prev = L.apl_layout_mode(_lib.LAYOUT_HOST)
    assert prev == _lib.LAYOUT_DEVICE
    assert L.apl_layout_mode(prev) == _lib.LAYOUT_HOST


def test_morton_order_arguments_checked_before_any_device_call():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = _lib.lib()
    assert L.apl_morton_order(None, 10, None, 0, None, None, 0, None) == -1
    assert "Morton" in L.apl_last_error().decode()
    assert L.apl_morton_order(None, 0, None, 0, None, None, 0, None) == 0  # nothing to order


def test_device_mesh_binding_methods_present():
    # the binding layer's methods (their calls need a GPU; their presence does not)
    from apple_amd import runtime

    for name in ("close", "eval", "region", "info", "layout", "lumped_volume", "set_schedule", "set_activation",
                 "mixed_activation", "ext_force"):
        assert callable(getattr(runtime.DeviceMesh, name, None)), name
    assert callable(runtime.layout_mode) and callable(runtime.morton_order)
