"""dpl_trace_noalloc, the trace that allocates nothing (its ray counter and traversal stacks live
in a fixed per-device scratch array), called through the C ABI: hits equal dpl_trace's bit for
bit on two calls in a row, the second with a smaller input, on trees whose spill stacks need
the persistent grid capped and on small ones."""
import ctypes as C
import sys

import numpy as np
import pytest

from diplodocus_amd import scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def noalloc(ctx, bvh, rays, flags, hits):
    from diplodocus_amd import _lib

    return _lib.lib.dpl_trace_noalloc(ctx.h, bvh.h, C.c_void_p(rays.data_ptr()), int(rays.shape[0]), flags,
                                      C.c_void_p(hits.data_ptr()))


def test_noalloc_trace_equals_trace(ctx):
    import torch

    from diplodocus_amd import _lib, accel

    big = scenes.mixed_2m()
    gb = accel.GpuBvh(ctx).build(torch.from_numpy(big.tris).cuda(),
                                 accel.make_params(morton_type=1, radius=16, max_leaf=4, bv_type=1))
    assert gb.stats.tree_height > 30  # the spill stacks of a full grid would not fit the scratch
    room = scenes.icosphere_room()
    gr = accel.GpuBvh(ctx).build(room.tris, accel.make_params(0, 16, 4))
    cases = [(gb, scenes.random_rays_towards(big.tris, 1 << 20, seed=5), 0),
             (gr, scenes.camera_rays(room.camera, 61, 47), 0),
             (gb, scenes.random_rays_towards(big.tris, 70_001, seed=6), _lib.DPL_TRACE_ANY),
             (gr, scenes.camera_rays(room.camera, 17, 9), _lib.DPL_TRACE_CULL_BACK)]
    for k, (g, rays_np, flags) in enumerate(cases):
        say(f"noalloc call {k}: {len(rays_np)} rays, flags {flags}")
        rays = torch.from_numpy(rays_np).cuda()
        hits = torch.full((len(rays_np), 4), -7.0, dtype=torch.float32, device="cuda")
        st = noalloc(ctx, g, rays, flags, hits)
        assert st == 0, _lib.lib.dpl_last_error(ctx.h).decode()
        ref = g.trace(rays, flags).clone()
        torch.cuda.synchronize()
        assert torch.equal(hits.view(torch.int32), ref.view(torch.int32)), k
    # what it does not take
    rays = torch.from_numpy(scenes.camera_rays(room.camera, 8, 8)).cuda()
    hits = torch.zeros((64, 4), dtype=torch.float32, device="cuda")
    assert noalloc(ctx, gr, rays, _lib.DPL_TRACE_WIDE, hits) == 1
    assert noalloc(ctx, gr, rays[:0], 0, hits) == 0
    gb.close()
    gr.close()
