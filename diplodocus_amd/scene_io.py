"""Scene loading from Python: ctypes binding of libdpl_scene.so (include/dpl/dpl_scene.h).

``load_scene(dirpath, name)`` is the reference's ObjSceneLoader::Load
(src/io/scene/obj_scene_loader.cc:188-204): <dirpath>/<name>.meta (camera, point lights) and
<dirpath>/<name>.obj with its material libraries, read by the engine's C++ loader
(diplodocus_amd/host/scene_loader.cc). The result holds numpy arrays in the layouts the GPU
entry points take: positions n x 9, vertex normals n x 9 (zero rows where a triangle has
none), material ids, dpl_material rows for the Whitted renderer, area lights in triangle
order, and the camera as a scenes.Camera.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import scenes

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdpl_scene.so")

EXPORTED = ["dpl_scene_load", "dpl_scene_load_obj", "dpl_scene_destroy", "dpl_scene_last_error",
            "dpl_scene_get_info", "dpl_scene_get_triangles", "dpl_scene_get_materials",
            "dpl_scene_get_material_name", "dpl_scene_get_area_lights", "dpl_scene_get_point_lights"]


class SceneInfo(C.Structure):
    _fields_ = [("triangle_count", C.c_int64), ("material_count", C.c_int32), ("area_light_count", C.c_int32),
                ("point_light_count", C.c_int32), ("reserved", C.c_int32), ("camera", C.c_float * 11)]


class SceneLoadError(RuntimeError):
    pass


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "(make -C diplodocus_amd/host)")
        L = C.CDLL(LIB_PATH)
        P, I32 = C.c_void_p, C.c_int32
        L.dpl_scene_load.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(P)]
        L.dpl_scene_load_obj.argtypes = [C.c_char_p, C.POINTER(P)]
        L.dpl_scene_destroy.argtypes = [P]
        L.dpl_scene_destroy.restype = None
        L.dpl_scene_last_error.argtypes = []
        L.dpl_scene_last_error.restype = C.c_char_p
        L.dpl_scene_get_info.argtypes = [P, C.POINTER(SceneInfo)]
        L.dpl_scene_get_triangles.argtypes = [P, P, P, P, P, P]
        L.dpl_scene_get_materials.argtypes = [P, P, P]
        L.dpl_scene_get_material_name.argtypes = [P, I32, C.c_char_p, I32]
        L.dpl_scene_get_area_lights.argtypes = [P, P, P, P]
        L.dpl_scene_get_point_lights.argtypes = [P, P]
        for name in EXPORTED:
            if name not in ("dpl_scene_destroy", "dpl_scene_last_error"):
                getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


@dataclass
class LoadedScene:
    tris: np.ndarray            # (n, 9) float32 vertex positions
    vertex_normals: np.ndarray  # (n, 9) float32, zero rows where a triangle has none
    has_vertex_normals: np.ndarray  # (n,) bool
    mat_id: np.ndarray          # (n,) int32, -1 = no material
    geom_normal: np.ndarray     # (n, 3) float32
    materials: np.ndarray       # (m, 19) float32: ambient3 diffuse3 specular3 transmittance3 emission3
                                #   shininess ior r_ior dissolve
    render_materials: np.ndarray  # (m, 16) float32, dpl_material rows
    material_names: list
    light_tri: np.ndarray       # (k,) int32 area light triangles, in triangle order
    light_color: np.ndarray     # (k, 3) float32
    light_area: np.ndarray      # (k,) float32
    point_lights: np.ndarray    # (p, 7) float32 pos3 color3 power
    camera_array: np.ndarray    # (11,) float32 pos3 dir3 up3 fov far

    @property
    def n(self) -> int:
        return len(self.tris)

    @property
    def camera(self) -> scenes.Camera:
        c = [float(x) for x in self.camera_array]
        return scenes.Camera(pos=tuple(c[0:3]), dir=tuple(c[3:6]), up=tuple(c[6:9]), fov=c[9], far=c[10])

    def vertex_normals_or_none(self):
        return self.vertex_normals if self.has_vertex_normals.any() else None


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _fetch(L, h) -> LoadedScene:
    info = SceneInfo()
    _check(L, L.dpl_scene_get_info(h, C.byref(info)))
    n, m, k, p = info.triangle_count, info.material_count, info.area_light_count, info.point_light_count
    tris = np.zeros((n, 9), np.float32)
    vn = np.zeros((n, 9), np.float32)
    has = np.zeros(n, np.uint8)
    mat = np.zeros(n, np.int32)
    gn = np.zeros((n, 3), np.float32)
    _check(L, L.dpl_scene_get_triangles(h, _ptr(tris), _ptr(vn), _ptr(has), _ptr(mat), _ptr(gn)))
    rows = np.zeros((m, 19), np.float32)
    render = np.zeros((m, 16), np.float32)
    _check(L, L.dpl_scene_get_materials(h, _ptr(rows), _ptr(render)))
    names = []
    buf = C.create_string_buffer(256)
    for i in range(m):
        _check(L, L.dpl_scene_get_material_name(h, i, buf, 256))
        names.append(buf.value.decode("latin-1"))
    lt = np.zeros(k, np.int32)
    lc = np.zeros((k, 3), np.float32)
    la = np.zeros(k, np.float32)
    _check(L, L.dpl_scene_get_area_lights(h, _ptr(lt), _ptr(lc), _ptr(la)))
    pl = np.zeros((p, 7), np.float32)
    _check(L, L.dpl_scene_get_point_lights(h, _ptr(pl)))
    return LoadedScene(tris, vn, has.astype(bool), mat, gn, rows, render, names, lt, lc, la, pl,
                       np.array(info.camera[:], np.float32))


def _check(L, status: int) -> None:
    if status != 0:
        raise SceneLoadError(L.dpl_scene_last_error().decode("latin-1", "replace"))


def _run(loader, *args) -> LoadedScene:
    L = lib()
    h = C.c_void_p()
    st = loader(*args, C.byref(h))
    if st != 0:
        raise SceneLoadError(L.dpl_scene_last_error().decode("latin-1", "replace"))
    try:
        return _fetch(L, h)
    finally:
        L.dpl_scene_destroy(h)


def load_scene(dirpath: str, name: str) -> LoadedScene:
    """ObjSceneLoader::Load of <dirpath>/<name>.{meta,obj}; raises SceneLoadError."""
    L = lib()
    return _run(L.dpl_scene_load, os.fsencode(dirpath), os.fsencode(name))


def load_obj(obj_path: str) -> LoadedScene:
    """ObjSceneLoader::LoadObj of one OBJ file (default camera, no point lights)."""
    L = lib()
    return _run(L.dpl_scene_load_obj, os.fsencode(obj_path))
