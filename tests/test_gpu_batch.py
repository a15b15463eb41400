"""GPU parity of the batched build and trace: every mesh of a batch equals the single-scene
build of that mesh (bit-exact against the oracle and against the single-scene GPU path),
batched traces equal per-mesh traces, captured trace graphs replay the eager result, and the
autotuner picks from its candidate set and serves repeats from its JSON cache."""
import json

import numpy as np
import pytest

from diplodocus_amd import scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def mixed_batch(seed, extra=40):
    """Mesh sizes that exercise single triangles, tiny meshes, the LDS-resident tail
    (<= 1024 clusters), chunked iterations (> 1024) over one and several chunks."""
    rng = np.random.default_rng(seed)
    sizes = [1, 2, 3, 5, 17, 64, 1000, 1024, 1025, 2049, 4096, 7000, 12000]
    sizes += [int(x) for x in rng.integers(1, 3000, extra)]
    parts = []
    for k, n in enumerate(sizes):
        c = rng.uniform(-5, 5, (n, 1, 3)) * rng.uniform(0.1, 3.0) + rng.uniform(-50, 50, 3)
        parts.append((c + rng.normal(0, 0.2, (n, 3, 3))).reshape(n, 9))
    off = np.zeros(len(parts) + 1, np.int64)
    off[1:] = np.cumsum([len(p) for p in parts])
    return np.vstack(parts).astype(np.float32), off


def mesh_slice(d, mi):
    a, k = int(mi["node_offset"]), int(mi["node_count"])
    p, n = int(mi["prim_offset"]), int(mi["prim_count"])
    return d["nodes"][a:a + k], d["prim_order"][p:p + n], d["tri_ordered"][12 * p:12 * (p + n)]


@pytest.mark.parametrize("mt,radius,leaf,block,digit", [(0, 16, 4, 256, 8), (1, 8, 4, 512, 4), (1, 32, 15, 256, 8),
                                                        (0, 1, 1, 512, 8)])
def test_batched_build_bitexact(oracle, ctx, mt, radius, leaf, block, digit):
    from diplodocus_amd import accel

    say(f"batched mt={mt} r={radius} leaf={leaf} block={block} digit={digit}")
    tris, off = mixed_batch(radius * 31 + leaf)
    p = accel.make_params(morton_type=mt, radius=radius, max_leaf=leaf, block_size=block, sort_digit_bits=digit)
    g = accel.GpuBvh(ctx).build_batched(tris, off, p)
    d = g.download()
    meshes = g.meshes()
    assert len(meshes) == len(off) - 1
    for m in range(len(off) - 1):
        tm = tris[off[m]:off[m + 1]]
        o = oracle.Ploc(tm, mt, radius, leaf)
        nodes, prim, tord = mesh_slice(d, meshes[m])
        assert len(nodes) == len(o.fnodes), m
        assert (nodes.view(np.uint32) == o.fnodes.view(np.uint32)).all(), f"final nodes of mesh {m}"
        assert (prim == o.prim_order).all(), f"prim order of mesh {m}"
        assert (tord.view(np.uint32) == o.tri_ordered.view(np.uint32)).all(), f"ordered triangles of mesh {m}"
        assert meshes[m]["ploc_iterations"] == len(o.iter_counts), m
        assert meshes[m]["sah_cost"] == pytest.approx(o.sah(), rel=1e-9), m
    g.close()


def test_batched_equals_single_scene_gpu_builds(ctx):
    import torch

    from diplodocus_amd import accel

    tris, off = scenes.batched_meshes(count=48, lo=64, hi=20000, seed=9)
    p = accel.make_params(morton_type=1, radius=16, max_leaf=4)
    say(f"batched vs single: {len(off) - 1} meshes, {off[-1]} tris")
    g = accel.GpuBvh(ctx).build_batched(torch.from_numpy(tris).cuda(), off, p)
    d = g.download()
    meshes = g.meshes()
    single = accel.GpuBvh(ctx)
    for m in range(len(off) - 1):
        s = single.build(np.ascontiguousarray(tris[off[m]:off[m + 1]]), p).download()
        nodes, prim, tord = mesh_slice(d, meshes[m])
        assert (nodes.view(np.uint32) == s["nodes"].view(np.uint32)).all(), m
        assert (prim == s["prim_order"]).all(), m
        assert (tord.view(np.uint32) == s["tri_ordered"].view(np.uint32)).all(), m
        assert meshes[m]["sah_cost"] == pytest.approx(single.stats.sah_cost, rel=1e-12), m
    assert g.stats.build_ms > 0
    g.close()
    single.close()


def batch_rays(tris, off, per_mesh, seed):
    rays, ids = [], []
    for m in range(len(off) - 1):
        rays.append(scenes.random_rays_towards(tris[off[m]:off[m + 1]], per_mesh, seed=seed + m))
        ids.append(np.full(per_mesh, m, np.int32))
    return np.vstack(rays), np.concatenate(ids)


def test_batched_trace_matches_per_mesh_traces(oracle, ctx):
    import torch

    from diplodocus_amd import accel

    tris, off = scenes.batched_meshes(count=40, lo=64, hi=8000, seed=4)
    p = accel.make_params(morton_type=1, radius=16, max_leaf=4)
    g = accel.GpuBvh(ctx).build_batched(tris, off, p)
    rays, ids = batch_rays(tris, off, 256, 100)
    perm = np.random.default_rng(1).permutation(len(ids))  # interleave meshes within waves
    rays, ids = np.ascontiguousarray(rays[perm]), np.ascontiguousarray(ids[perm])
    say(f"batched trace: {len(rays)} rays over {len(off) - 1} meshes")
    h_host = g.trace_batched(rays, ids)
    h_dev = g.trace_batched(torch.from_numpy(rays).cuda(), torch.from_numpy(ids).cuda()).cpu().numpy()
    assert (h_dev.view(np.uint32).reshape(-1, 4) == h_host.view(np.uint32).reshape(-1, 4)).all()
    single = accel.GpuBvh(ctx)
    for m in range(len(off) - 1):
        sel = ids == m
        tm = np.ascontiguousarray(tris[off[m]:off[m + 1]])
        ref = single.build(tm, p).trace(np.ascontiguousarray(rays[sel]))
        assert (h_host[sel].view(np.uint32) == ref.view(np.uint32)).all(), m
        if m % 8 == 0:
            bf = oracle.trace_brute(tm, np.ascontiguousarray(rays[sel]))
            assert (h_host[sel]["prim"] == bf["prim"]).all(), m
    # mesh ids out of range give misses
    bad = ids.copy()
    bad[::7] = len(off) - 1
    bad[3::7] = -1
    hb = g.trace_batched(rays, bad)
    assert (hb["prim"][(bad < 0) | (bad >= len(off) - 1)] == -1).all()
    ok = (bad >= 0) & (bad < len(off) - 1)
    assert (hb[ok].view(np.uint32) == h_host[ok].view(np.uint32)).all()
    single.close()
    g.close()


def test_trace_graph_replays_match_eager(ctx):
    import torch

    from diplodocus_amd import accel

    tris, off = scenes.batched_meshes(count=24, lo=64, hi=4000, seed=8)
    g = accel.GpuBvh(ctx).build_batched(tris, off, accel.make_params(morton_type=1, radius=16, max_leaf=4))
    rays_np, ids_np = batch_rays(tris, off, 256, 7)
    rays = torch.from_numpy(rays_np).cuda()
    ids = torch.from_numpy(ids_np).cuda()
    hits_graph = torch.zeros((len(rays_np), 4), dtype=torch.float32, device="cuda")
    graph = accel.TraceGraph(g, rays, hits_graph, ray_mesh=ids)
    one = accel.GpuBvh(ctx).build(scenes.icosphere_room().tris, accel.make_params(0, 16, 4))
    cam = torch.from_numpy(scenes.camera_rays(scenes.Camera(), 64, 64)).cuda()
    cam_hits = torch.zeros((cam.shape[0], 4), dtype=torch.float32, device="cuda")
    graph1 = accel.TraceGraph(one, cam, cam_hits)
    for k in range(4):
        say(f"graph replay {k}")
        fresh, _ = batch_rays(tris, off, 256, 1000 * (k + 1))
        rays.copy_(torch.from_numpy(fresh))
        eager = g.trace_batched(rays, ids).clone()  # eager work between replays
        graph.launch()
        graph1.launch()
        torch.cuda.synchronize()
        assert torch.equal(hits_graph.view(torch.int32), eager.view(torch.int32)), k
        assert torch.equal(cam_hits.view(torch.int32), one.trace(cam).view(torch.int32)), k
    graph.close()
    graph1.close()
    one.close()
    g.close()


def test_trace_graph_invalidated_by_rebuild_widen_destroy(ctx):
    """A graph keeps the BVH's device addresses and sizes: after a rebuild, a widen or a
    destroy of its BVH, launching it is refused without running anything."""
    import torch

    from diplodocus_amd import _lib, accel

    tris = scenes.icosphere_room().tris
    g = accel.GpuBvh(ctx).build(tris, accel.make_params(0, 16, 4))
    cam = torch.from_numpy(scenes.camera_rays(scenes.Camera(), 32, 32)).cuda()
    hits = torch.zeros((cam.shape[0], 4), dtype=torch.float32, device="cuda")
    gr = accel.TraceGraph(g, cam, hits)
    gr.launch()
    torch.cuda.synchronize()
    assert torch.equal(hits.view(torch.int32), g.trace(cam).view(torch.int32))
    say("graph after rebuild")
    g.build(tris[: len(tris) // 2].copy(), accel.make_params(0, 16, 4))  # rebuild in place, smaller
    with pytest.raises(accel.DplError) as e:
        gr.launch()
    assert e.value.status == 1 and "rebuilt" in str(e.value)
    gr2 = accel.TraceGraph(g, cam, hits)  # a new capture of the new tree replays
    gr2.launch()
    torch.cuda.synchronize()
    assert torch.equal(hits.view(torch.int32), g.trace(cam).view(torch.int32))
    g.widen(4)
    with pytest.raises(accel.DplError):
        gr2.launch()
    gr3 = accel.TraceGraph(g, cam, hits, flags=_lib.DPL_TRACE_WIDE)
    gr3.launch()
    torch.cuda.synchronize()
    g.close()
    with pytest.raises(accel.DplError):
        gr3.launch()
    for x in (gr, gr2, gr3):
        x.close()


def test_batched_error_paths(ctx):
    from diplodocus_amd import accel

    tris, off = mixed_batch(3, extra=2)
    g = accel.GpuBvh(ctx)
    cases = [(dict(bv_type=1), off, 4), (dict(block_size=128), off, 1), (dict(), off[::-1].copy(), 1),
             (dict(), off + 1, 1), (dict(), np.array([0, 5, 5, len(tris)], np.int64), 1)]
    for kw, o, status in cases:
        with pytest.raises(accel.DplError) as e:
            g.build_batched(tris, o, accel.make_params(**kw))
        assert e.value.status == status, (kw, status)
    bad = tris.copy()
    bad[off[5] + 1, 3] = np.inf
    with pytest.raises(accel.DplError) as e:
        g.build_batched(bad, off, accel.make_params())
    assert e.value.status == 1
    g.build_batched(tris, off, accel.make_params())
    rays = scenes.random_rays_towards(tris, 64, seed=1)
    with pytest.raises(accel.DplError) as e:
        g.trace(rays)
    assert e.value.status == 1
    g.close()


def test_autotune_picks_a_candidate_and_caches(ctx, tmp_path):
    from diplodocus_amd import accel

    cache = str(tmp_path / "autotune.json")
    tris = scenes.icosphere_room().tris
    base = accel.make_params(morton_type=1, radius=16, max_leaf=4)
    best, info = accel.autotune(ctx, tris, base, sah_tolerance=0.05, cache_path=cache)
    say(f"autotune single: r={best.nn_search_radius} block={best.block_size} digit={best.sort_digit_bits} {info}")
    assert best.nn_search_radius in (8, 16, 32) and best.block_size in (128, 256, 512)
    assert best.sort_digit_bits in (4, 8) and best.morton_type == 1 and best.max_triangles_per_leaf == 4
    assert info["candidates"] == 18 and not info["from_cache"] and info["build_ms"] > 0
    again, info2 = accel.autotune(ctx, tris, base, cache_path=cache)
    assert info2["from_cache"] and info2["candidates"] == 0
    assert (again.nn_search_radius, again.block_size, again.sort_digit_bits) == \
           (best.nn_search_radius, best.block_size, best.sort_digit_bits)
    # the chosen build is within the SAH tolerance of every candidate's best
    g = accel.GpuBvh(ctx).build(tris, best)
    assert g.stats.sah_cost == pytest.approx(info["sah_cost"], rel=1e-12)
    g.close()
    btris, boff = scenes.batched_meshes(count=32, lo=64, hi=4000, seed=2)
    bbest, binfo = accel.autotune(ctx, btris, base, offsets=boff, cache_path=cache)
    assert bbest.block_size in (256, 512) and binfo["candidates"] == 12
    entries = json.load(open(cache))["entries"]
    assert len(entries) == 2 and any(k.startswith("batched/") for k in entries)


def test_full_batch_4096_sampled_meshes_bitexact(oracle, ctx):
    """The bench's whole batch (4096 meshes, log-uniform 64 to 65,536 triangles, ~38M) in one
    launch with its autotuned choice (PLOC+EMC, r = 8, 512-thread blocks, 8-bit digits): every
    64th mesh and the eight largest equal the oracle's build of that mesh alone bit for bit, and
    the graph-captured trace of 256 rays per mesh gives those meshes the oracle's hits."""
    import torch

    from diplodocus_amd import accel

    tris, off = scenes.batched_meshes()
    counts = np.diff(off)
    say(f"full batch: {len(off) - 1} meshes, {int(off[-1])} triangles")
    dev = torch.device("cuda:0")
    p = accel.make_params(morton_type=1, radius=8, max_leaf=4, block_size=512, sort_digit_bits=8)
    g = accel.GpuBvh(ctx).build_batched(torch.from_numpy(tris).to(dev), off, p)
    d = g.download()
    meshes = g.meshes()
    rays, ids = scenes.batched_rays(tris, off, 256)
    hits = torch.empty((len(ids), 4), dtype=torch.float32, device=dev)
    graph = accel.TraceGraph(g, torch.from_numpy(rays).to(dev), hits, ray_mesh=torch.from_numpy(ids).to(dev))
    graph.launch()
    ctx.synchronize()
    h = hits.cpu().numpy().view(np.uint32).reshape(-1, 4)
    sample = sorted(set(range(0, len(counts), 64)) | set(np.argsort(counts)[-8:].tolist()))
    for m in sample:
        tm = np.ascontiguousarray(tris[off[m]:off[m + 1]])
        o = oracle.Ploc(tm, 1, 8, 4)
        nodes, prim, tord = mesh_slice(d, meshes[m])
        assert (nodes.view(np.uint32) == o.fnodes.view(np.uint32)).all(), f"final nodes of mesh {m}"
        assert (prim == o.prim_order).all(), f"prim order of mesh {m}"
        assert (tord.view(np.uint32) == o.tri_ordered.view(np.uint32)).all(), f"ordered triangles of mesh {m}"
        sel = np.nonzero(ids == m)[0]
        ref = oracle.trace_final(o.fnodes, o.tri_ordered, np.ascontiguousarray(rays[sel]))
        assert (h[sel] == ref.view(np.uint32).reshape(-1, 4)).all(), f"hits of mesh {m}"
    say(f"  {len(sample)} meshes equal (largest {int(counts.max())} triangles)")
    graph.close()
    g.close()
