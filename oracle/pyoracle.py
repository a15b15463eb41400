"""ctypes wrapper of the oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker
and the timed CPU baseline; never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

PNODE_DTYPE = np.dtype([("lo", "<f4", 3), ("hi", "<f4", 3), ("left", "<i4"), ("right", "<i4")])
FNODE_DTYPE = np.dtype([("box", "<f4", 12), ("left", "<i4"), ("right", "<i4"), ("pad", "<i4", 2)])
HIT_DTYPE = np.dtype([("t", "<f4"), ("b1", "<f4"), ("b2", "<f4"), ("prim", "<i4")])


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = C.CDLL(_LIB)
    P, I32, I64, U32, U64, F, D = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_float, C.c_double
    sig = {
        "orc_hash_u32": (U32, [U32]),
        "orc_u01": (F, [U32]),
        "orc_ray_epsilon": (F, [F, F, F, F]),
        "orc_centroid": (None, [P, P]),
        "orc_triangle_area": (F, [P]),
        "orc_intersect_mt": (F, [P, P, I32, P, P]),
        "orc_intersect_watertight": (I32, [P, P, I32, P, P, P]),
        "orc_intersect_box": (I32, [P, P, P, F, P]),
        "orc_intersect_box_ref": (I32, [P, P, P, P, P]),
        "orc_intersect_fp64": (D, [P, P, P, P]),
        "orc_sbvh_build": (P, [P, I64, I32, I32, I32, I32, P, P]),
        "orc_sbvh_free": (None, [P]),
        "orc_sbvh_node_capacity": (I64, [P]),
        "orc_sbvh_nodes_used": (I64, [P]),
        "orc_sbvh_export": (None, [P, P, P, P]),
        "orc_sbvh_sah": (D, [P]),
        "orc_sbvh_trace": (D, [P, P, I64, U32, I32, P, I32, P]),
        "orc_bounds": (None, [P, I64, P]),
        "orc_expand_bits10": (U32, [U32]),
        "orc_codes": (None, [P, I64, I32, P]),
        "orc_emc_layout": (None, [P, P, P]),
        "orc_ploc_nn": (None, [P, I32, I32, P]),
        "orc_sa_union": (F, [P, P, P, P]),
        "orc_ploc_build": (P, [P, I64, I32, I32, I32, I32, I32, P]),
        "orc_ploc_kdop": (I32, [P, P]),
        "orc_sobb_set": (I32, [I32, P, P, P, P]),
        "orc_ploc_free": (None, [P]),
        "orc_ploc_iterations": (I32, [P]),
        "orc_ploc_final_nodes": (I64, [P]),
        "orc_ploc_export": (None, [P] * 14),
        "orc_ploc_sah": (D, [P]),
        "orc_ploc_trace": (None, [P, P, I64, U32, P, I32]),
        "orc_ploc_trace_counted": (None, [P, P, I64, U32, P, I32, P]),
        "orc_trace_final": (None, [P, I64, P, I64, P, I64, U32, P, I32, I32]),
        "orc_trace_final_counted": (None, [P, I64, P, I64, P, I64, U32, P, I32, I32, P]),
        "orc_sah_final": (D, [P, I64, P, I32]),
        "orc_refit_aabb": (I32, [P, I64, P, I64, P]),
        "orc_widen": (I32, [P, I64, I32, P]),
        "orc_wide_node_bytes": (I32, [I32]),
        "orc_trace_wide": (None, [P, I64, I32, P, I64, P, I64, U32, P, I32, P]),
        "orc_trace_brute": (None, [P, I64, P, I64, U32, P, I32]),
        "orc_trace_dummy": (None, [P, I64, P, I64, U32, P, I32]),
        "orc_pcg_stream": (None, [U64, U64, I32, P]),
        "orc_sincos2pi": (None, [F, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def nthreads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


# ---- small helpers -----------------------------------------------------------------------
def hash_u32(x: int) -> int:
    return int(lib.orc_hash_u32(x))


def centroid(tri9) -> np.ndarray:
    t = _f32(tri9)
    out = np.zeros(3, np.float32)
    lib.orc_centroid(_p(t), _p(out))
    return out


def intersect_mt(tri9, ray8, cull=False):
    t, r = _f32(tri9), _f32(ray8)
    b1, b2 = C.c_float(), C.c_float()
    tt = lib.orc_intersect_mt(_p(t), _p(r), int(cull), C.byref(b1), C.byref(b2))
    return tt, b1.value, b2.value


def intersect_watertight(tri9, ray8, cull=False):
    t, r = _f32(tri9), _f32(ray8)
    tt, b1, b2 = C.c_float(), C.c_float(), C.c_float()
    ok = lib.orc_intersect_watertight(_p(t), _p(r), int(cull), C.byref(tt), C.byref(b1), C.byref(b2))
    return bool(ok), tt.value, b1.value, b2.value


def intersect_fp64(tri9, ray8):
    t, r = _f32(tri9), _f32(ray8)
    b1, b2 = C.c_double(), C.c_double()
    tt = lib.orc_intersect_fp64(_p(t), _p(r), C.byref(b1), C.byref(b2))
    return tt, b1.value, b2.value


def bounds(tris) -> np.ndarray:
    t = _f32(tris)
    out = np.zeros(12, np.float32)
    lib.orc_bounds(_p(t), t.shape[0], _p(out))
    return out


def codes(tris, morton_type: int) -> np.ndarray:
    t = _f32(tris)
    out = np.zeros(t.shape[0], np.uint64)
    lib.orc_codes(_p(t), t.shape[0], morton_type, _p(out))
    return out


def emc_layout(bounds12):
    b = _f32(bounds12)
    nb = np.zeros(3, np.int32)
    ax = np.zeros(48, np.int8)
    lib.orc_emc_layout(_p(b), _p(nb), _p(ax))
    return nb, ax


def ploc_nn(lohi, radius: int) -> np.ndarray:
    b = _f32(lohi)
    out = np.zeros(b.shape[0], np.int32)
    lib.orc_ploc_nn(_p(b), b.shape[0], radius, _p(out))
    return out


def trace_brute(tris, rays, flags=0, threads=None) -> np.ndarray:
    t, r = _f32(tris), _f32(rays)
    hits = np.zeros(r.shape[0], HIT_DTYPE)
    lib.orc_trace_brute(_p(t), t.shape[0], _p(r), r.shape[0], flags, _p(hits), threads or nthreads())
    return hits


def trace_dummy(tris, rays, flags=0, threads=None) -> np.ndarray:
    """The reference's brute force (dummy_acceleration_structure.cc:21-79): MT with the
    absolute |det| < 1e-6 rejection, closest hit strict (first of equal t wins), any-hit at
    t <= t_max. prim = input index, -1 on a miss; any-hit results carry the hit found."""
    t, r = _f32(tris), _f32(rays)
    hits = np.zeros(r.shape[0], HIT_DTYPE)
    lib.orc_trace_dummy(_p(t), t.shape[0], _p(r), r.shape[0], flags, _p(hits), threads or nthreads())
    return hits


def trace_final(fnodes, tri_ordered, rays, flags=0, threads=None, sobb_kdop=0) -> np.ndarray:
    """Traces a tree in the engine's final layout (e.g. downloaded from the device);
    sobb_kdop: 0 for AABB child boxes, else the kdop_size of the SOBB direction set."""
    fn = np.ascontiguousarray(fnodes)
    to = _f32(tri_ordered)
    r = _f32(rays)
    hits = np.zeros(r.shape[0], HIT_DTYPE)
    lib.orc_trace_final(_p(fn), fn.shape[0], _p(to), to.size // 12, _p(r), r.shape[0], flags, _p(hits),
                        threads or nthreads(), sobb_kdop)
    return hits


def trace_final_counted(fnodes, tri_ordered, rays, flags=0, threads=None, sobb_kdop=0):
    """trace_final plus (inner nodes visited, triangle tests) summed over the rays."""
    fn = np.ascontiguousarray(fnodes)
    to = _f32(tri_ordered)
    r = _f32(rays)
    hits = np.zeros(r.shape[0], HIT_DTYPE)
    counts = np.zeros(2, np.int64)
    lib.orc_trace_final_counted(_p(fn), fn.shape[0], _p(to), to.size // 12, _p(r), r.shape[0], flags, _p(hits),
                                threads or nthreads(), sobb_kdop, _p(counts))
    return hits, int(counts[0]), int(counts[1])


def refit_aabb(fnodes, tri_ordered) -> np.ndarray:
    """Axis-aligned child boxes of every node of a final tree (same references, zero pad)."""
    fn = np.ascontiguousarray(fnodes)
    to = _f32(tri_ordered)
    out = np.zeros(fn.shape[0], FNODE_DTYPE)
    if lib.orc_refit_aabb(_p(fn), fn.shape[0], _p(to), to.size // 12, _p(out)) != 0:
        raise ValueError("tree too deep")
    return out


def wide_node_bytes(width: int) -> int:
    return int(lib.orc_wide_node_bytes(width))


def widen(aabb_fnodes, width: int) -> np.ndarray:
    """Wide nodes (uint8 rows of wide_node_bytes(width)) from AABB final nodes."""
    fn = np.ascontiguousarray(aabb_fnodes)
    out = np.zeros((fn.shape[0], wide_node_bytes(width)), np.uint8)
    if lib.orc_widen(_p(fn), fn.shape[0], width, _p(out)) != 0:
        raise ValueError("width must be 4 or 8")
    return out


def trace_wide(wnodes, width, tri_ordered, rays, flags=0, threads=None):
    """(hits, wide nodes visited, triangle tests) over a wide tree."""
    wn = np.ascontiguousarray(wnodes, dtype=np.uint8)
    to = _f32(tri_ordered)
    r = _f32(rays)
    hits = np.zeros(r.shape[0], HIT_DTYPE)
    counts = np.zeros(2, np.int64)
    lib.orc_trace_wide(_p(wn), wn.shape[0], width, _p(to), to.size // 12, _p(r), r.shape[0], flags, _p(hits),
                       threads or nthreads(), _p(counts))
    return hits, int(counts[0]), int(counts[1])


def sah_final(fnodes, root_box6, sobb_kdop=0) -> float:
    fn = np.ascontiguousarray(fnodes)
    rb = _f32(root_box6)
    return float(lib.orc_sah_final(_p(fn), fn.shape[0], _p(rb), sobb_kdop))


def sobb_set(kdop_size=32):
    nd = C.c_int32()
    dirs = np.zeros((13, 3), np.float32)
    abc = np.zeros((286, 3), np.int32)
    inv = np.zeros(286, np.float32)
    nt = lib.orc_sobb_set(kdop_size, C.byref(nd), _p(dirs), _p(abc), _p(inv))
    return dirs[: nd.value], abc[:nt], inv[:nt]


def pcg_stream(seed: int, stream: int, count: int) -> np.ndarray:
    out = np.zeros(count, np.uint32)
    lib.orc_pcg_stream(seed, stream, count, _p(out))
    return out


def sincos2pi(u: float):
    s, c = C.c_float(), C.c_float()
    lib.orc_sincos2pi(u, C.byref(s), C.byref(c))
    return s.value, c.value


# ---- SBvh --------------------------------------------------------------------------------
class SBvh:
    """Restated reference SAH BVH (src/acceleration/sbvh.cc)."""

    def __init__(self, tris, max_depth=32, max_leaf=4, bin_count=16, fix_child_bounds=False):
        self.tris = _f32(tris)
        st = np.zeros(3, np.int32)
        ms = C.c_double()
        self.h = lib.orc_sbvh_build(_p(self.tris), self.tris.shape[0], max_depth, max_leaf, bin_count,
                                    int(fix_child_bounds), _p(st), C.byref(ms))
        if not self.h:
            raise ValueError("empty scene")
        self.node_count, self.inner_node_count, self.leaf_node_count = (int(x) for x in st)
        self.build_ms = ms.value

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_sbvh_free(self.h)
            self.h = None

    def export(self):
        cap = lib.orc_sbvh_node_capacity(self.h)
        mm = np.zeros((cap, 6), np.float32)
        bc = np.zeros((cap, 2), np.int32)
        ti = np.zeros(self.tris.shape[0], np.int32)
        lib.orc_sbvh_export(self.h, _p(mm), _p(bc), _p(ti))
        used = lib.orc_sbvh_nodes_used(self.h)
        return mm[:used], bc[:used], ti

    def sah(self) -> float:
        return float(lib.orc_sbvh_sah(self.h))

    def trace(self, rays, flags=0, mode=1, threads=None):
        r = _f32(rays)
        hits = np.zeros(r.shape[0], HIT_DTYPE)
        vis = C.c_int64()
        secs = lib.orc_sbvh_trace(self.h, _p(r), r.shape[0], flags, mode, _p(hits), threads or nthreads(),
                                  C.byref(vis))
        self.last_trace_s = secs
        self.last_visits = vis.value
        return hits


# ---- PLOC restatement --------------------------------------------------------------------
class Ploc:
    """Sequential restatement of the engine's PLOC build (codes, order, merges, collapse,
    final layout)."""

    def __init__(self, tris, morton_type=0, radius=16, max_leaf=4, bv_type=0, kdop_size=32):
        self.tris = _f32(tris)
        ms = C.c_double()
        self.h = lib.orc_ploc_build(_p(self.tris), self.tris.shape[0], morton_type, radius, max_leaf, bv_type,
                                    kdop_size, C.byref(ms))
        if not self.h:
            raise ValueError("empty scene")
        self.build_ms = ms.value
        self.sobb_kdop = kdop_size if bv_type == 1 else 0
        n = self.tris.shape[0]
        self.n = n
        nd = lib.orc_ploc_kdop(self.h, None)
        self.kdop = None
        if nd:
            self.kdop = np.zeros((2 * n - 1, 2 * nd), np.float32)
            lib.orc_ploc_kdop(self.h, _p(self.kdop))
        it = lib.orc_ploc_iterations(self.h)
        nf = lib.orc_ploc_final_nodes(self.h)
        self.codes = np.zeros(n, np.uint64)
        self.sorted_codes = np.zeros(n, np.uint64)
        self.sorted_idx = np.zeros(n, np.uint32)
        self.nodes = np.zeros(2 * n - 1, PNODE_DTYPE)
        self.parent = np.zeros(2 * n - 1, np.int32)
        self.iter_counts = np.zeros(it, np.int32)
        self.cost = np.zeros(2 * n - 1, np.float32)
        self.count = np.zeros(2 * n - 1, np.int32)
        self.collapsed = np.zeros(2 * n - 1, np.uint8)
        self.fnodes = np.zeros(nf, FNODE_DTYPE)
        self.prim_order = np.zeros(n, np.int32)
        self.tri_ordered = np.zeros(12 * n, np.float32)
        self.bounds = np.zeros(12, np.float32)
        arrs = [self.codes, self.sorted_codes, self.sorted_idx, self.nodes, self.parent, self.iter_counts, self.cost,
                self.count, self.collapsed, self.fnodes, self.prim_order, self.tri_ordered, self.bounds]
        lib.orc_ploc_export(self.h, *[_p(a) for a in arrs])

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_ploc_free(self.h)
            self.h = None

    def sah(self) -> float:
        return float(lib.orc_ploc_sah(self.h))

    def trace(self, rays, flags=0, threads=None):
        r = _f32(rays)
        hits = np.zeros(r.shape[0], HIT_DTYPE)
        lib.orc_ploc_trace(self.h, _p(r), r.shape[0], flags, _p(hits), threads or nthreads())
        return hits

    def trace_counted(self, rays, flags=0, threads=None):
        """(hits, inner nodes visited, triangle tests) over this tree."""
        r = _f32(rays)
        hits = np.zeros(r.shape[0], HIT_DTYPE)
        counts = np.zeros(2, np.int64)
        lib.orc_ploc_trace_counted(self.h, _p(r), r.shape[0], flags, _p(hits), threads or nthreads(), _p(counts))
        return hits, int(counts[0]), int(counts[1])


# ---- diffuse path tracing ------------------------------------------------------------------
def _sig_render():
    P, I32, I64, U64, D = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double
    for name, res, args in [
        ("orc_camera_basis", None, [P, I32, I32, P]),
        ("orc_render", None, [P, P, P, P, P, I32, I32, I32, I32, U64, P, P, P, I32]),
        ("orc_path_loss_fp64", D, [P, I64, I32, I32, P, P, P]),
        ("orc_path_grad_fp64", None, [P, I64, I32, I32, I32, P, P, P, P, P]),
    ]:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


_sig_render()


def render(ploc: "Ploc", scene, W, H, spp, depth, seed, albedo=None, emission=None, record_paths=False,
           threads=None):
    """CPU restatement of the engine's diffuse path tracer over the oracle's PLOC tree.
    Returns (image (H,W,3) float32, path materials (W*H*spp, depth) int32 or None)."""
    alb = _f32(scene.albedo if albedo is None else albedo)
    emi = _f32(scene.emission if emission is None else emission)
    mat = np.ascontiguousarray(scene.mat_id, np.int32)
    cam = scene.camera.as_array()
    img = np.zeros((H, W, 3), np.float32)
    pm = np.full((W * H * spp, depth), -1, np.int32) if record_paths else None
    lib.orc_render(ploc.h, _p(ploc.tris), _p(mat), _p(alb), _p(emi), W, H, spp, depth, seed, _p(cam), _p(img),
                   _p(pm) if pm is not None else None, threads or nthreads())
    return img, pm


def path_loss_fp64(pm, pixels, spp, depth, albedo64, emission64, target) -> float:
    a = np.ascontiguousarray(albedo64, np.float64)
    e = np.ascontiguousarray(emission64, np.float64)
    t = _f32(target)
    return float(lib.orc_path_loss_fp64(_p(pm), pixels, spp, depth, _p(a), _p(e), _p(t)))


def path_grad_fp64(pm, pixels, spp, depth, albedo64, emission64, target):
    a = np.ascontiguousarray(albedo64, np.float64)
    e = np.ascontiguousarray(emission64, np.float64)
    t = _f32(target)
    ga = np.zeros_like(a)
    ge = np.zeros_like(e)
    lib.orc_path_grad_fp64(_p(pm), pixels, spp, depth, a.shape[0], _p(a), _p(e), _p(t), _p(ga), _p(ge))
    return ga, ge


# ---- Whitted ray tracing --------------------------------------------------------------------
def _sig_whitted():
    P, I32, I64, U32, F = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_float
    for name, res, args in [
        ("orc_whitted", None, [P, P, P, I64, P, P, I32, I32, I32, I32, I32, P, P, I32, P, I32, P]),
        ("orc_pow_pos", F, [F, F]),
        ("orc_area_light_sample01", F, [U32, U32, U32, U32, U32, U32]),
    ]:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


_sig_whitted()


def pow_pos(x: float, y: float) -> float:
    return lib.orc_pow_pos(x, y)


def area_light_sample01(seed, px, py, light, s, dim) -> float:
    return lib.orc_area_light_sample01(seed, px, py, light, s, dim)


WHITTED_COUNTS = ("primary_ray_count", "secondary_ray_count", "shadow_ray_count", "query_count",
                  "traversal_count", "intersection_count")


def whitted(ploc: "Ploc", mat_id, materials, W, H, max_depth, samples, camera, background=(0.0, 0.0, 0.0),
            cull=False, vertex_normals=None, recursive=False, threads=None, counts=False):
    """CPU restatement of the reference's Whitted CpuRaytracer over the oracle's PLOC tree
    (watertight hits). recursive=True evaluates the reference's recursion as written;
    otherwise the weighted depth-first sum the GPU kernel uses. Returns (H, W, 3) float32, and
    with counts=True also the reference's ray counters and the traversal work as a dict keyed
    by WHITTED_COUNTS (traversal_count = inner nodes visited, intersection_count = triangle
    tests)."""
    mats = _f32(materials).reshape(-1, 16)
    mat = np.ascontiguousarray(mat_id, np.int32)
    cam = _f32(camera.as_array() if hasattr(camera, "as_array") else camera)
    bg = _f32(background)
    vn = _f32(vertex_normals) if vertex_normals is not None else None
    img = np.zeros((H, W, 3), np.float32)
    c = np.zeros(6, np.int64)
    lib.orc_whitted(ploc.h, _p(ploc.tris), _p(vn) if vn is not None else None, ploc.tris.shape[0], _p(mat), _p(mats),
                    W, H, max_depth, samples, 1 if cull else 0, _p(bg), _p(cam), 1 if recursive else 0, _p(img),
                    threads or nthreads(), _p(c) if counts else None)
    if counts:
        return img, dict(zip(WHITTED_COUNTS, (int(v) for v in c)))
    return img
