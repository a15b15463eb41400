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
from ._lib import RenderParams, lib


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
