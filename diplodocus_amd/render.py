"""Differentiable diffuse path tracing over a GpuBvh, and a thin torch.autograd.Function.

``render_forward`` / ``render_adjoint`` call dpl_render_forward / dpl_render_adjoint (the
engine's replacement of the reference's LaunchPathtracingKernel,
src/gpu/renderer/gpu_renderer_api.h:11). ``RenderFunction`` wraps them so that the image-loss
gradient with respect to per-material albedo and emission comes from the adjoint path-replay
kernel; it exists for checking the adjoint against finite differences.
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import RenderParams, WhittedParams, lib


def make_render_params(width: int, height: int, spp: int, max_depth: int, seed: int, camera,
                       material_count: int) -> RenderParams:
    rp = RenderParams()
    rp.width, rp.height, rp.spp, rp.max_depth = width, height, spp, max_depth
    rp.seed = seed
    cam = camera.as_array() if hasattr(camera, "as_array") else camera
    for k in range(11):
        rp.camera[k] = float(cam[k])
    rp.material_count = material_count
    return rp


def _dev(t):
    import torch

    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous()):
        raise ValueError("render buffers must be contiguous GPU tensors")
    return C.c_void_p(t.data_ptr())


def render_forward(bvh, rp: RenderParams, mat_id, albedo, emission, image=None):
    import torch

    if image is None:
        image = torch.empty((rp.height, rp.width, 3), dtype=torch.float32, device=albedo.device)
    bvh.ctx.check(lib.dpl_render_forward(bvh.ctx.h, bvh.h, C.byref(rp), _dev(mat_id), _dev(albedo), _dev(emission),
                                         _dev(image)))
    return image


def render_adjoint(bvh, rp: RenderParams, mat_id, albedo, emission, d_image):
    import torch

    ga = torch.empty_like(albedo)
    ge = torch.empty_like(emission)
    bvh.ctx.check(lib.dpl_render_adjoint(bvh.ctx.h, bvh.h, C.byref(rp), _dev(mat_id), _dev(albedo), _dev(emission),
                                         _dev(d_image.contiguous()), _dev(ga), _dev(ge)))
    return ga, ge


def render_function():
    """Returns a torch.autograd.Function: image = F(albedo (M,3), emission (M,3))."""
    import torch

    class RenderFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, albedo, emission, bvh, rp, mat_id):
            albedo = albedo.contiguous()
            emission = emission.contiguous()
            img = render_forward(bvh, rp, mat_id, albedo, emission)
            ctx.bvh, ctx.rp, ctx.mat_id = bvh, rp, mat_id
            ctx.save_for_backward(albedo, emission)
            return img

        @staticmethod
        def backward(ctx, grad_img):
            albedo, emission = ctx.saved_tensors
            ga, ge = render_adjoint(ctx.bvh, ctx.rp, ctx.mat_id, albedo, emission, grad_img)
            return ga, ge, None, None, None

    return RenderFunction


# ---- Whitted ray tracing (dpl_render_whitted) -----------------------------------------------
def make_whitted_params(width: int, height: int, max_depth: int, samples: int, camera, background=(0.0, 0.0, 0.0),
                        backface_culling: bool = False) -> WhittedParams:
    wp = WhittedParams()
    wp.width, wp.height, wp.max_depth, wp.area_light_sample_cnt = width, height, max_depth, samples
    wp.backface_culling = 1 if backface_culling else 0
    for k in range(3):
        wp.background[k] = float(background[k])
    cam = camera.as_array() if hasattr(camera, "as_array") else camera
    for k in range(11):
        wp.camera[k] = float(cam[k])
    return wp


def emissive_triangles(mat_id, materials) -> "np.ndarray":
    """The reference's area lights: triangles whose material emits, in triangle order
    (src/io/scene/obj_scene_loader.cc:95-105)."""
    import numpy as np

    mats = np.asarray(materials, np.float32).reshape(-1, 16)
    emits = (mats[:, 9:12] != 0).any(axis=1)
    return np.nonzero(emits[np.asarray(mat_id)])[0].astype(np.int32)


def render_whitted(bvh, wp: WhittedParams, tris, mat_id, materials, lights, vertex_normals=None, image=None,
                   stats=False):
    """GPU Whitted image (H, W, 3). tris / mat_id / materials (M, 16) / lights / vertex_normals
    are GPU tensors (materials rows: diffuse3 specular3 transmittance3 emission3 shininess ior
    r_ior pad). stats=True (dpl_render_whitted_stats) returns (image, counters dict): the
    reference's ray counters, the queries with their traversal work, and the kernel time."""
    import torch

    if image is None:
        image = torch.empty((wp.height, wp.width, 3), dtype=torch.float32, device=tris.device)
    m = int(materials.shape[0])
    nl = int(lights.shape[0])
    args = (bvh.ctx.h, bvh.h, C.byref(wp), _dev(tris), _dev(vertex_normals) if vertex_normals is not None else None,
            _dev(mat_id), _dev(materials), m, _dev(lights) if nl else None, nl, _dev(image))
    if stats:
        st = _lib.WhittedStats()
        bvh.ctx.check(lib.dpl_render_whitted_stats(*args, C.byref(st)))
        return image, st.as_dict()
    bvh.ctx.check(lib.dpl_render_whitted(*args))
    return image
