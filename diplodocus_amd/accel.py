"""Host-side mirror of the reference's acceleration interface over the C ABI.

Names follow the reference: ``AccelerationStructureType`` and ``AccelerationStructureConfig``
(src/config/acceleration_structure_config.h:9-29), ``bridge_acceleration_config``
(src/gpu/config/gpu_config_bridge.cc:24-60) and an ``AccelerationStructure``-shaped
``GpuBvh`` with ``build`` / ``intersect`` / ``intersect_any``
(src/acceleration/acceleration_structure.h:12-19), batched over ray arrays.

Arrays may be numpy arrays (host memory) or torch tensors on the GPU (device memory, used
in place). Every call goes to libdpl; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import BuildParams, BuildStats, BvhLayout, DplError, lib

HIT_DTYPE = np.dtype([("t", "<f4"), ("b1", "<f4"), ("b2", "<f4"), ("prim", "<i4")])
PNODE_DTYPE = np.dtype([("lo", "<f4", 3), ("hi", "<f4", 3), ("left", "<i4"), ("right", "<i4")])
FNODE_DTYPE = np.dtype([("box", "<f4", 12), ("left", "<i4"), ("right", "<i4"), ("pad", "<i4", 2)])
MESH_DTYPE = np.dtype([("node_offset", "<i8"), ("prim_offset", "<i8"), ("node_count", "<i4"), ("prim_count", "<i4"),
                       ("tree_height", "<i4"), ("ploc_iterations", "<i4"), ("sah_cost", "<f8")])


class AccelerationStructureType(enum.IntEnum):
    """src/config/acceleration_structure_config.h:9"""

    kDummy = 0
    kSbvh = 1
    kPloc = 2
    kPlocEmc = 3
    kPlocSobb = 4
    kPlocEmcSobb = 5

    @classmethod
    def parse(cls, s: str) -> "AccelerationStructureType":
        """ParseEnum<AccelerationStructureType> (acceleration_structure_config.h:11-20)."""
        table = {"dummy": cls.kDummy, "sbvh": cls.kSbvh, "ploc": cls.kPloc, "ploc_emc": cls.kPlocEmc,
                 "ploc_sobb": cls.kPlocSobb, "ploc_emc_sobb": cls.kPlocEmcSobb}
        if s not in table:
            raise RuntimeError("Invalid AccelerationStructureType: " + s)
        return table[s]


@dataclass
class AccelerationStructureConfig:
    """src/config/acceleration_structure_config.h:22-29 (same defaults)."""

    acceleration_structure_type: AccelerationStructureType = AccelerationStructureType.kDummy
    max_depth: int = 32
    nn_search_radius: int = 10
    kdop_size: int = 32
    max_triangles_per_leaf: int = 4
    bin_count: int = 16


def bridge_acceleration_config(cfg: AccelerationStructureConfig, block_size: int = 0, sort_digit_bits: int = 0,
                               keep_intermediates: bool = False) -> BuildParams:
    """Maps a PLOC type to (Morton30 | EMC60) x (AABB | SOBB) plus radius / leaf / k-DOP
    (src/gpu/config/gpu_config_bridge.cc:24-60). Dummy and SBvh are CPU structures."""
    p = BuildParams()
    st = lib.dpl_build_params_from_type(int(cfg.acceleration_structure_type), cfg.nn_search_radius,
                                        cfg.max_triangles_per_leaf, cfg.kdop_size, C.byref(p))
    if st != _lib.DPL_OK:
        raise DplError(st, f"no GPU build for {cfg.acceleration_structure_type!r}")
    p.block_size = block_size
    p.sort_digit_bits = sort_digit_bits
    p.keep_intermediates = 1 if keep_intermediates else 0
    return p


def make_params(morton_type: int = 0, radius: int = 16, max_leaf: int = 4, bv_type: int = 0, block_size: int = 0,
                sort_digit_bits: int = 0, keep_intermediates: int = 0, kdop_size: int = 32) -> BuildParams:
    """keep_intermediates: bit 0 keeps codes/order/merge list, bit 1 times every PLOC
    iteration launch (True = 1)."""
    return BuildParams(morton_type, bv_type, radius, max_leaf, kdop_size, block_size, sort_digit_bits,
                       int(keep_intermediates))


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


def _ptr(a):
    """(pointer, mem kind) of a numpy array (host) or a torch CUDA tensor (device)."""
    if _is_torch(a):
        if not a.is_cuda:
            raise ValueError("torch tensors must live on the GPU; pass numpy arrays for host memory")
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return C.c_void_p(a.data_ptr()), _lib.DPL_MEM_DEVICE
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(C.c_void_p), _lib.DPL_MEM_HOST


class Context:
    """A dpl context: one device, one stream, scratch pools."""

    def __init__(self, device: int = 0, use_torch_stream: bool = True):
        h = C.c_void_p()
        st = lib.dpl_context_create(device, C.byref(h))
        if st != _lib.DPL_OK:
            raise DplError(st, "dpl_context_create failed (no GPU visible?)")
        self.h = h
        self.device = device
        if use_torch_stream:
            try:
                import torch

                if torch.cuda.is_available():
                    self.set_stream(torch.cuda.current_stream(device).cuda_stream)
            except ImportError:
                pass

    def check(self, st: int) -> None:
        if st != _lib.DPL_OK:
            raise DplError(st, lib.dpl_last_error(self.h).decode())

    def set_stream(self, stream: int) -> None:
        self.check(lib.dpl_context_set_stream(self.h, C.c_void_p(stream)))

    def synchronize(self) -> None:
        self.check(lib.dpl_context_synchronize(self.h))

    @property
    def cu_count(self) -> int:
        return int(lib.dpl_context_cu_count(self.h))

    def close(self) -> None:
        if getattr(self, "h", None):
            lib.dpl_context_destroy(self.h)
            self.h = None

    def __del__(self):
        # at interpreter exit the module globals (lib) may already be cleared; the
        # process's device memory goes with it
        if lib is not None:
            self.close()

    # ---- stand-alone stages (parity checks / autotuner) ----
    def compute_codes(self, tris_dev, morton_type: int, codes_dev):
        """Device tensors: tris (n,9) float32, codes (n,) int64 (uint64 bits). Returns the
        12 bounds floats."""
        b = np.zeros(12, np.float32)
        tp, mem = _ptr(tris_dev)
        cp, _ = _ptr(codes_dev)
        if mem != _lib.DPL_MEM_DEVICE:
            raise ValueError("compute_codes takes device tensors")
        self.check(lib.dpl_compute_codes(self.h, tp, tris_dev.shape[0], morton_type, cp,
                                         b.ctypes.data_as(C.c_void_p)))
        return b

    def radix_sort_pairs(self, keys_dev, vals_dev, key_bits: int, digit_bits: int = 8) -> None:
        kp, mem = _ptr(keys_dev)
        vp, _ = _ptr(vals_dev)
        if mem != _lib.DPL_MEM_DEVICE:
            raise ValueError("radix_sort_pairs takes device tensors")
        self.check(lib.dpl_radix_sort_pairs(self.h, kp, vp, keys_dev.shape[0], key_bits, digit_bits))


class GpuBvh:
    """A PLOC BVH on the device; the GPU counterpart of SBvh behind the same interface."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.h = None
        self.stats = None

    # AccelerationStructure::Build (acceleration_structure.h:15-16)
    def build(self, tris, params: BuildParams) -> "GpuBvh":
        tp, mem = _ptr(tris)
        n = int(tris.shape[0])
        st = BuildStats()
        if self.h is None:
            h = C.c_void_p()
            self.ctx.check(lib.dpl_bvh_build(self.ctx.h, tp, n, mem, C.byref(params), C.byref(h), C.byref(st)))
            self.h = h
        else:
            self.ctx.check(lib.dpl_bvh_rebuild(self.h, tp, n, mem, C.byref(params), C.byref(st)))
        self.stats = st
        self.n = n
        return self

    def build_batched(self, tris, offsets, params: BuildParams) -> "GpuBvh":
        """One BVH per mesh tris[offsets[m]:offsets[m+1]], all built in one persistent launch."""
        tp, mem = _ptr(tris)
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        st = BuildStats()
        self.close()
        h = C.c_void_p()
        self.ctx.check(lib.dpl_bvh_build_batched(self.ctx.h, tp, off.ctypes.data_as(C.c_void_p), len(off) - 1, mem,
                                                 C.byref(params), C.byref(h), C.byref(st)))
        self.h = h
        self.stats = st
        self.n = int(off[-1])
        return self

    def meshes(self) -> np.ndarray:
        """Per-mesh node/triangle offsets and counts, tree height, iterations and SAH."""
        lay = self.layout()
        out = (_lib.MeshInfo * int(lay.mesh_count))()
        self.ctx.check(lib.dpl_bvh_get_meshes(self.h, out))
        return np.array([(m.node_offset, m.prim_offset, m.node_count, m.prim_count, m.tree_height, m.ploc_iterations,
                          m.sah_cost) for m in out], MESH_DTYPE)

    def trace_batched(self, rays, ray_mesh, flags: int = 0, hits=None):
        """Ray i against mesh ray_mesh[i] of a batched BVH; hit prim ids are mesh-local."""
        rp, mem = _ptr(rays)
        mp, mmem = _ptr(ray_mesh)
        n = int(rays.shape[0])
        if hits is None:
            if mem == _lib.DPL_MEM_DEVICE:
                import torch

                hits = torch.empty((n, 4), dtype=torch.float32, device=rays.device)
            else:
                hits = np.zeros(n, HIT_DTYPE)
        hp, hmem = _ptr(hits)
        if not (hmem == mem == mmem):
            raise ValueError("rays, ray_mesh and hits must live in the same memory")
        self.ctx.check(lib.dpl_trace_batched(self.ctx.h, self.h, rp, mp, n, flags, hp, mem))
        return hits

    def close(self) -> None:
        if self.h is not None:
            lib.dpl_bvh_destroy(self.h)
            self.h = None

    def __del__(self):
        # at interpreter exit the module globals (lib) may already be cleared; the
        # process's device memory goes with it
        if lib is not None:
            self.close()

    def sah_aabb(self) -> float:
        """SAH cost with axis-aligned child boxes (a SOBB tree's boxes refitted); equals
        stats.sah_cost for AABB trees."""
        v = C.c_double()
        self.ctx.check(lib.dpl_bvh_sah_aabb(self.h, C.byref(v)))
        return v.value

    def layout(self) -> BvhLayout:
        lay = BvhLayout()
        self.ctx.check(lib.dpl_bvh_get_layout(self.h, C.byref(lay)))
        return lay

    # AccelerationStructure::Intersect / IntersectAny (acceleration_structure.h:17-18),
    # over a batch of 32-byte rays.
    def trace(self, rays, flags: int = 0, hits=None):
        rp, mem = _ptr(rays)
        n = int(rays.shape[0])
        if hits is None:
            if mem == _lib.DPL_MEM_DEVICE:
                import torch

                hits = torch.empty((n, 4), dtype=torch.float32, device=rays.device)
            else:
                hits = np.zeros(n, HIT_DTYPE)
        hp, hmem = _ptr(hits)
        if hmem != mem:
            raise ValueError("rays and hits must live in the same memory")
        self.ctx.check(lib.dpl_trace(self.ctx.h, self.h, rp, n, flags, hp, mem))
        return hits

    def trace_counters(self, rays, flags: int = 0):
        """Trace plus (inner nodes visited, triangle tests) over the batch."""
        rp, mem = _ptr(rays)
        n = int(rays.shape[0])
        if mem == _lib.DPL_MEM_DEVICE:
            import torch

            hits = torch.empty((n, 4), dtype=torch.float32, device=rays.device)
        else:
            hits = np.zeros(n, HIT_DTYPE)
        hp, _ = _ptr(hits)
        nv, nt = C.c_int64(), C.c_int64()
        self.ctx.check(lib.dpl_trace_counters(self.ctx.h, self.h, rp, n, flags, hp, mem, C.byref(nv), C.byref(nt)))
        return hits, int(nv.value), int(nt.value)

    def intersect(self, rays, backface_culling: bool = False):
        return self.trace(rays, _lib.DPL_TRACE_CULL_BACK if backface_culling else 0)

    def intersect_any(self, rays, backface_culling: bool = False):
        return self.trace(rays, _lib.DPL_TRACE_ANY | (_lib.DPL_TRACE_CULL_BACK if backface_culling else 0))

    def download(self, intermediates: bool = False) -> dict:
        lay = self.layout()
        n, k = int(lay.prim_count), int(lay.node_count)
        out = {"nodes": np.zeros(k, FNODE_DTYPE), "tri_ordered": np.zeros(12 * n, np.float32),
               "prim_order": np.zeros(n, np.int32)}
        ptrs = [out["nodes"].ctypes.data_as(C.c_void_p), out["tri_ordered"].ctypes.data_as(C.c_void_p),
                out["prim_order"].ctypes.data_as(C.c_void_p)]
        if intermediates:
            out["codes"] = np.zeros(n, np.uint64)
            out["sorted_idx"] = np.zeros(n, np.uint32)
            out["ploc_nodes"] = np.zeros(int(lay.ploc_node_count), PNODE_DTYPE)
            out["iter_counts"] = np.zeros(int(lay.iteration_count), np.int32)
            ptrs += [out[k].ctypes.data_as(C.c_void_p) for k in ("codes", "sorted_idx", "ploc_nodes", "iter_counts")]
        else:
            ptrs += [None] * 4
        self.ctx.check(lib.dpl_bvh_download(self.h, *ptrs))
        return out

    # ---- wide tree (no counterpart in the reference): 4 or 8 children, quantized boxes ----
    def widen(self, width: int = 4) -> dict:
        """Adds the wide form; trace(..., flags | DPL_TRACE_WIDE) then traverses it."""
        info = _lib.WideInfo()
        self.ctx.check(lib.dpl_bvh_widen(self.h, width, C.byref(info)))
        return {k: getattr(info, k) for k, _ in info._fields_}

    def wide_info(self) -> dict:
        info = _lib.WideInfo()
        self.ctx.check(lib.dpl_bvh_get_wide(self.h, C.byref(info)))
        return {k: getattr(info, k) for k, _ in info._fields_}

    def download_wide(self) -> np.ndarray:
        """Wide nodes as uint8 rows of node_bytes."""
        info = self.wide_info()
        out = np.zeros((int(info["node_count"]), int(info["node_bytes"])), np.uint8)
        self.ctx.check(lib.dpl_bvh_download_wide(self.h, out.ctypes.data_as(C.c_void_p)))
        return out


class TraceGraph:
    """A trace pass captured once in a hipGraph and replayed by launch(); rays, ray_mesh
    (batched BVHs only) and hits are GPU tensors whose storage the graph keeps using."""

    def __init__(self, bvh: GpuBvh, rays, hits, ray_mesh=None, flags: int = 0):
        rp, mem = _ptr(rays)
        hp, hmem = _ptr(hits)
        mp = None
        if ray_mesh is not None:
            mp, mmem = _ptr(ray_mesh)
            if mmem != _lib.DPL_MEM_DEVICE:
                raise ValueError("ray_mesh must be a GPU tensor")
        if mem != _lib.DPL_MEM_DEVICE or hmem != _lib.DPL_MEM_DEVICE:
            raise ValueError("a captured trace needs GPU tensors")
        self.ctx, self._keep = bvh.ctx, (bvh, rays, hits, ray_mesh)
        h = C.c_void_p()
        self.ctx.check(lib.dpl_trace_graph_create(self.ctx.h, bvh.h, rp, mp, int(rays.shape[0]), flags, hp,
                                                  C.byref(h)))
        self.h = h

    def launch(self) -> None:
        self.ctx.check(lib.dpl_trace_graph_launch(self.h))

    def close(self) -> None:
        if getattr(self, "h", None) is not None:
            lib.dpl_trace_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        # at interpreter exit the module globals (lib) may already be cleared; the
        # process's device memory goes with it
        if lib is not None:
            self.close()


def autotune(ctx: Context, tris, base: BuildParams, offsets=None, sah_tolerance: float = 0.05,
             cache_path: str | None = None):
    """Chooses nn_search_radius, block_size and sort_digit_bits for this workload (see
    dpl_autotune); returns (params, info dict)."""
    tp, mem = _ptr(tris)
    best = BuildParams()
    info = _lib.AutotuneInfo()
    if offsets is not None:
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        op, m, n = off.ctypes.data_as(C.c_void_p), len(off) - 1, int(off[-1])
    else:
        off, op, m, n = None, None, 0, int(tris.shape[0])
    ctx.check(lib.dpl_autotune(ctx.h, tp, n, op, m, mem, C.byref(base), sah_tolerance,
                               cache_path.encode() if cache_path else None, C.byref(best), C.byref(info)))
    return best, {"build_ms": info.build_ms, "sah_cost": info.sah_cost, "candidates": info.candidates,
                  "from_cache": bool(info.from_cache)}
