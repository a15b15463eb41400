"""The C-ABI libraries load and export every symbol their headers declare (include/dpl/dpl.h:
libdpl.so; include/dpl/dpl_scene.h: libdpl_scene.so); host-side logic that needs no GPU
(parameter bridge, error paths)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols(header="dpl.h"):
    text = open(os.path.join(ROOT, "include", "dpl", header)).read()
    decl = r"^(?:dpl_status|void|const char\s*\*|int32_t)\s+(dpl_[a-z_0-9]+)\s*\("
    return sorted(set(re.findall(decl, text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    from diplodocus_amd import _lib

    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(_lib.lib, s), s
    assert sorted(_lib.EXPORTED) == syms


def test_scene_library_exports_every_declared_symbol():
    from diplodocus_amd import scene_io

    syms = header_symbols("dpl_scene.h")
    assert len(syms) == 10
    for s in syms:
        assert hasattr(scene_io.lib(), s), s
    assert sorted(scene_io.EXPORTED) == syms


def test_scene_library_null_arguments():
    from diplodocus_amd import scene_io

    L = scene_io.lib()
    assert L.dpl_scene_load(None, b"x", None) == 1
    assert b"NULL" in L.dpl_scene_last_error()
    assert L.dpl_scene_load_obj(None, None) == 1
    assert L.dpl_scene_get_info(None, None) == 1
    assert L.dpl_scene_get_triangles(None, None, None, None, None, None) == 1
    assert L.dpl_scene_get_material_name(None, 0, None, 0) == 1
    L.dpl_scene_destroy(None)


def test_params_bridge_mirrors_reference_switch():
    # src/gpu/config/gpu_config_bridge.cc:27-53
    from diplodocus_amd import accel

    T = accel.AccelerationStructureType
    expect = {T.kPloc: (0, 0), T.kPlocEmc: (1, 0), T.kPlocSobb: (0, 1), T.kPlocEmcSobb: (1, 1)}
    for t, (m, b) in expect.items():
        cfg = accel.AccelerationStructureConfig(t, nn_search_radius=8, max_triangles_per_leaf=3, kdop_size=26)
        p = accel.bridge_acceleration_config(cfg)
        assert (p.morton_type, p.bv_type, p.nn_search_radius, p.max_triangles_per_leaf, p.kdop_size) == (m, b, 8, 3, 26)
    for t in (T.kDummy, T.kSbvh):
        with pytest.raises(accel.DplError):
            accel.bridge_acceleration_config(accel.AccelerationStructureConfig(t))
    assert accel.AccelerationStructureType.parse("ploc_emc_sobb") == T.kPlocEmcSobb
    with pytest.raises(RuntimeError):
        accel.AccelerationStructureType.parse("bvh8")


def test_null_and_invalid_arguments_without_device():
    from diplodocus_amd import _lib

    L = _lib.lib
    assert L.dpl_build_params_from_type(2, 16, 4, 32, None) == _lib.DPL_ERR_INVALID_ARG
    p = _lib.BuildParams()
    assert L.dpl_build_params_from_type(99, 16, 4, 32, C.byref(p)) == _lib.DPL_ERR_INVALID_ARG
    assert L.dpl_context_create(0, None) == _lib.DPL_ERR_INVALID_ARG
    assert L.dpl_context_synchronize(None) == _lib.DPL_ERR_INVALID_ARG
    assert L.dpl_bvh_get_stats(None, None) == _lib.DPL_ERR_INVALID_ARG
    assert L.dpl_trace(None, None, None, 0, 0, None, 0) == _lib.DPL_ERR_INVALID_ARG
    assert L.dpl_last_error(None) == b"null context"
    L.dpl_context_destroy(None)
    L.dpl_bvh_destroy(None)


def test_context_create_reports_missing_device():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from diplodocus_amd import _lib

    h = C.c_void_p()
    assert _lib.lib.dpl_context_create(0, C.byref(h)) == _lib.DPL_ERR_DEVICE
    assert not h.value


def test_destructors_tolerate_interpreter_exit(monkeypatch):
    """At interpreter exit the module global holding the library can be cleared before the
    last handles are collected: the destructors then do nothing instead of raising."""
    from diplodocus_amd import accel

    objs = [cls.__new__(cls) for cls in (accel.Context, accel.GpuBvh, accel.TraceGraph)]
    for o in objs:
        o.h = 12345  # never passed to the library: lib is cleared below
    monkeypatch.setattr(accel, "lib", None)
    for o in objs:
        o.__del__()
        o.h = None
