"""GPU parity of the wide trees: the converted nodes are byte-identical to the oracle's
conversion of the same two-child tree (SOBB trees through the axis-aligned refit); the wide
trace gives the oracle's wide traversal bit for bit, counters included, and the two-child
trace's closest hits; a captured graph replays it; a rebuild drops the wide form."""
import sys

import numpy as np
import pytest

from diplodocus_amd import scenes

from .test_wide import SCENES, rays_for

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_bvh(ctx, tris, mt=1, radius=8, leaf=4, bv=0):
    from diplodocus_amd import accel

    return accel.GpuBvh(ctx).build(tris, accel.make_params(morton_type=mt, radius=radius, max_leaf=leaf, bv_type=bv))


def same(a, b):
    return a.view(np.uint32).tobytes() == b.view(np.uint32).tobytes()


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("bv", [0, 1])
@pytest.mark.parametrize("width", [4, 8])
def test_gpu_widen_matches_oracle(oracle, ctx, name, bv, width):
    say(f"widen {name} bv={bv} width={width}")
    g = gpu_bvh(ctx, SCENES[name](), bv=bv)
    info = g.widen(width)
    d = g.download()
    assert info["width"] == width and info["node_count"] == len(d["nodes"])
    assert info["node_bytes"] == oracle.wide_node_bytes(width)
    w = g.download_wide()
    ref = oracle.widen(oracle.refit_aabb(d["nodes"], d["tri_ordered"]), width)
    assert w.tobytes() == ref.tobytes()


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("bv", [0, 1])
@pytest.mark.parametrize("width", [4, 8])
def test_gpu_wide_trace_matches_oracle(oracle, ctx, name, bv, width):
    from diplodocus_amd import _lib

    tris = SCENES[name]()
    say(f"wide trace {name} bv={bv} width={width}")
    g = gpu_bvh(ctx, tris, bv=bv)
    g.widen(width)
    d = g.download()
    w = g.download_wide()
    rays = rays_for(tris, 4000, seed=3 * width + bv)
    for flags in (0, 1, 2):
        h, nodes, tests = g.trace_counters(rays, flags | _lib.DPL_TRACE_WIDE)
        ref, rn, rt = oracle.trace_wide(w, width, d["tri_ordered"], rays, flags)
        assert same(h, ref) and (nodes, tests) == (rn, rt)
        if not flags & 1:
            assert same(h, g.trace(rays, flags))


@pytest.mark.parametrize("width", [4, 8])
def test_gpu_wide_trace_mixed_scene(oracle, ctx, width):
    """A 220k-triangle scene of clusters and thin triangles: wide closest hits equal the
    two-child trace's and brute force; wide nodes visited stay below two-child nodes."""
    from diplodocus_amd import _lib

    s = scenes.mixed_2m(n_small=200_000, n_thin=20_000)
    say(f"wide mixed width={width}")
    for bv in (0, 1):
        g = gpu_bvh(ctx, s.tris, radius=16, bv=bv)
        g.widen(width)
        rays = scenes.random_rays_towards(s.tris, 20_000, seed=width)
        hw, nw, _ = g.trace_counters(rays, _lib.DPL_TRACE_WIDE)
        hb, nb, _ = g.trace_counters(rays, 0)
        assert same(hw, hb) and nw < nb
        if bv == 1:
            sub = slice(0, 3000)
            assert same(hw[sub], oracle.trace_brute(s.tris, rays[sub]))


def test_gpu_wide_graph_replays(ctx):
    import torch

    from diplodocus_amd import _lib, accel

    s = scenes.mixed_2m(n_small=50_000, n_thin=5_000)
    say("wide graph")
    g = gpu_bvh(ctx, s.tris, bv=1)
    g.widen(4)
    rays = torch.from_numpy(scenes.random_rays_towards(s.tris, 30_000, seed=2)).cuda()
    hits = torch.zeros((30_000, 4), dtype=torch.float32, device="cuda")
    eager = g.trace(rays, _lib.DPL_TRACE_WIDE).cpu()
    gr = accel.TraceGraph(g, rays, hits, flags=_lib.DPL_TRACE_WIDE)
    for k in range(3):
        hits.zero_()
        torch.cuda.synchronize()
        gr.launch()
        ctx.synchronize()
        assert torch.equal(hits.cpu().view(torch.int32), eager.view(torch.int32)), k
    gr.close()


def test_gpu_rebuild_drops_wide(ctx):
    from diplodocus_amd import _lib, accel

    tris = SCENES["soup"]()
    say("rebuild drops wide")
    g = gpu_bvh(ctx, tris)
    g.widen(8)
    assert g.wide_info()["width"] == 8
    g.build(tris, accel.make_params(morton_type=1, radius=8, max_leaf=4))
    assert g.wide_info()["width"] == 0
    with pytest.raises(_lib.DplError):
        g.trace(rays_for(tris, 10, 1), _lib.DPL_TRACE_WIDE)
    with pytest.raises(_lib.DplError):
        g.widen(6)


def batch_of_meshes(count=48, seed=9):
    """Meshes of 1 to 3000 triangles (soups and flat patches), concatenated."""
    rng = np.random.default_rng(seed)
    sizes = [1, 2, 3] + [int(x) for x in np.exp(rng.uniform(np.log(4), np.log(3000), count - 3))]
    parts = [SCENES["flat"]()[: max(1, s)] if k % 5 == 4 else
             (rng.uniform(-1, 1, (s, 1, 3)) + rng.normal(0, 0.05, (s, 3, 3))).reshape(s, 9).astype(np.float32)
             for k, s in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum([len(p) for p in parts])]).astype(np.int64)
    return np.concatenate(parts).astype(np.float32), off


@pytest.mark.parametrize("width", [4, 8])
def test_gpu_batched_wide(oracle, ctx, width):
    """Batched BVHs: every mesh's wide nodes equal the oracle's conversion of that mesh's
    nodes (unused slots zero); wide hits equal the two-child batched trace's; a captured graph
    replays the wide batched trace."""
    import torch

    from diplodocus_amd import _lib, accel

    tris, off = batch_of_meshes()
    say(f"batched wide width={width} meshes={len(off) - 1}")
    g = accel.GpuBvh(ctx).build_batched(tris, off, accel.make_params(morton_type=1, radius=8, max_leaf=4,
                                                                     block_size=256))
    info = g.widen(width)
    d = g.download()
    ms = g.meshes()
    w = g.download_wide()
    assert info["node_count"] == len(d["nodes"]) == len(w)
    for m in range(len(ms)):
        a, k = int(ms[m]["node_offset"]), int(ms[m]["node_count"])
        end = int(ms[m + 1]["node_offset"]) if m + 1 < len(ms) else len(w)
        ref = oracle.widen(d["nodes"][a:a + k], width)
        assert w[a:a + k].tobytes() == ref.tobytes(), m
        assert not w[a + k:end].any(), m
    rays, ray_mesh = scenes.batched_rays(tris, off, 64, seed=width)
    for flags in (0, 1, 2):
        hw = g.trace_batched(rays, ray_mesh, flags | _lib.DPL_TRACE_WIDE)
        hb = g.trace_batched(rays, ray_mesh, flags)
        if flags & 1:
            assert ((hw["prim"] >= 0) == (hb["prim"] >= 0)).all()
        else:
            assert same(hw, hb)
    rd = torch.from_numpy(rays).cuda()
    md = torch.from_numpy(ray_mesh).cuda()
    hits = torch.zeros((len(rays), 4), dtype=torch.float32, device="cuda")
    eager = g.trace_batched(rd, md, _lib.DPL_TRACE_WIDE).cpu()
    gr = accel.TraceGraph(g, rd, hits, ray_mesh=md, flags=_lib.DPL_TRACE_WIDE)
    for k in range(3):
        hits.zero_()
        torch.cuda.synchronize()
        gr.launch()
        ctx.synchronize()
        assert torch.equal(hits.cpu().view(torch.int32), eager.view(torch.int32)), k
    gr.close()


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("kdop", [32, 14])
def test_sah_in_axis_aligned_areas(oracle, ctx, name, kdop):
    """dpl_bvh_sah_aabb: a SOBB tree's SAH with its children's axis-aligned boxes equals the
    oracle's SAH of the oracle's refit of the same tree (FP64 sums in another order: 1e-9); an
    AABB tree's is its own SAH."""
    from diplodocus_amd import accel

    tris = SCENES[name]()
    say(f"sah aabb {name} kdop={kdop}")
    g = accel.GpuBvh(ctx).build(tris, accel.make_params(morton_type=1, radius=16, max_leaf=4, bv_type=1,
                                                        kdop_size=kdop))
    d = g.download()
    v = tris.reshape(-1, 3)
    root6 = np.concatenate([v.min(0), v.max(0)]).astype(np.float32)
    ref = oracle.sah_final(oracle.refit_aabb(d["nodes"], d["tri_ordered"]), root6, 0)
    assert g.sah_aabb() == pytest.approx(ref, rel=1e-9)
    assert g.stats.sah_cost == pytest.approx(oracle.sah_final(d["nodes"], root6, kdop), rel=1e-9)
    a = gpu_bvh(ctx, tris, radius=16)
    assert a.sah_aabb() == a.stats.sah_cost
    g.close()
    a.close()
