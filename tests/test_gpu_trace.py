"""GPU parity of the trace: closest-hit results are bit-exact against the oracle traversal
of the same tree and against brute force; t agrees with an FP64 intersection to 1e-5;
any-hit and backface culling; the reference's own rules are compared and classified."""
import sys

import numpy as np
import pytest

from diplodocus_amd import scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def soup(n, seed, spread=10.0, size=0.2):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, spread, (n, 1, 3))
    return (c + rng.normal(0, size, (n, 3, 3))).reshape(n, 9).astype(np.float32)


def gpu_bvh(ctx, tris, mt=0, radius=16, leaf=4, bv=0, kdop=32):
    from diplodocus_amd import accel

    return accel.GpuBvh(ctx).build(tris, accel.make_params(morton_type=mt, radius=radius, max_leaf=leaf, bv_type=bv,
                                                           kdop_size=kdop))


@pytest.mark.parametrize("kdop", [32, 18, 14])
@pytest.mark.parametrize("flags", [0, 1, 2])
def test_sobb_trace_matches_brute_force(oracle, ctx, kdop, flags):
    s = scenes.mixed_2m(n_small=40_000, n_thin=6_000)
    say(f"sobb trace kdop={kdop} flags={flags}")
    g = gpu_bvh(ctx, s.tris, mt=1, bv=1, kdop=kdop)
    rays = scenes.random_rays_towards(s.tris, 8192, seed=kdop + flags)
    h = g.trace(rays, flags)
    brute = oracle.trace_brute(s.tris, rays, flags)
    if flags & 1:
        assert ((h["prim"] >= 0) == (brute["prim"] >= 0)).all()
    else:
        assert hits_equal(h, brute).all()
    d = g.download()
    # any-hit included: the first hit of the ordered traversal, whatever the wave around it
    assert hits_equal(h, oracle.trace_final(d["nodes"], d["tri_ordered"], rays, flags, sobb_kdop=kdop)).all()
    # SOBB trees visit fewer nodes than AABB trees on this scene of thin triangles
    _, va, _ = gpu_bvh(ctx, s.tris, mt=1).trace_counters(rays)
    _, vs, _ = g.trace_counters(rays)
    assert vs < va


def hits_equal(a, b):
    return (a.view(np.uint32).reshape(-1, 4) == b.view(np.uint32).reshape(-1, 4)).all(axis=1)


@pytest.mark.parametrize("bv,kdop", [(0, 32), (1, 32), (1, 14)])
@pytest.mark.parametrize("flags", [0, 1, 2])
def test_trace_counters_match_oracle(oracle, ctx, bv, kdop, flags):
    """The traversal counters (AccelerationStats::traversal_count / intersection_count as inner
    nodes visited and triangle tests) equal the oracle's count of the same ordered traversal
    over the same tree, ray for ray summed."""
    s = scenes.mixed_2m(n_small=30_000, n_thin=3_000)
    say(f"trace counters bv={bv} kdop={kdop} flags={flags}")
    g = gpu_bvh(ctx, s.tris, mt=1, bv=bv, kdop=kdop)
    rays = scenes.random_rays_towards(s.tris, 6000, seed=7 + flags)
    h, nodes, tests = g.trace_counters(rays, flags)
    d = g.download()
    ref, rn, rt = oracle.trace_final_counted(d["nodes"], d["tri_ordered"], rays, flags, sobb_kdop=kdop if bv else 0)
    assert (nodes, tests) == (rn, rt)
    assert hits_equal(h, ref).all()
    assert tests >= int((h["prim"] >= 0).sum()) and nodes >= len(rays)


@pytest.mark.parametrize("scene", ["cornell", "icosphere"])
@pytest.mark.parametrize("flags", [0, 2])
def test_trace_matches_brute_force(oracle, ctx, scene, flags):
    s = scenes.cornell_room() if scene == "cornell" else scenes.icosphere_room()
    say(f"trace {scene} flags={flags}")
    g = gpu_bvh(ctx, s.tris)
    rays = scenes.camera_rays(s.camera, 128, 128)
    h = g.trace(rays, flags)
    brute = oracle.trace_brute(s.tris, rays, flags)
    assert hits_equal(h, brute).all()
    d = g.download()
    assert hits_equal(h, oracle.trace_final(d["nodes"], d["tri_ordered"], rays, flags)).all()


@pytest.mark.parametrize("n,leaf", [(1, 4), (3, 4), (777, 1), (20000, 15)])
def test_trace_random_soups(oracle, ctx, n, leaf):
    tris = soup(n, n)
    g = gpu_bvh(ctx, tris, leaf=leaf)
    rays = scenes.random_rays_towards(tris, 4096, seed=n)
    assert hits_equal(g.trace(rays), oracle.trace_brute(tris, rays)).all()


def test_any_hit_agrees_with_closest(oracle, ctx):
    s = scenes.icosphere_room()
    g = gpu_bvh(ctx, s.tris)
    rays = scenes.camera_rays(s.camera, 96, 96)
    anyh = g.trace(rays, 1)
    closest = oracle.trace_brute(s.tris, rays)
    assert ((anyh["prim"] >= 0) == (closest["prim"] >= 0)).all()
    hit = anyh["prim"] >= 0
    assert (anyh["t"][hit] > rays[hit, 3]).all() and (anyh["t"][hit] < rays[hit, 7]).all()


def test_trace_t_against_fp64_and_device_buffers(oracle, ctx):
    import torch

    s = scenes.mixed_2m(n_small=150_000, n_thin=20_000)
    say("fp64 check on 170k-tri mixed scene")
    g = gpu_bvh(ctx, s.tris, mt=1)
    rays = scenes.random_rays_towards(s.tris, 1 << 18, seed=2)
    rd = torch.from_numpy(rays).cuda()
    hd = g.trace(rd)
    ctx.synchronize()
    h = hd.cpu().numpy().view(oracle.HIT_DTYPE).reshape(-1)
    host = g.trace(rays)
    assert hits_equal(h, host).all()
    idx = np.nonzero(h["prim"] >= 0)[0][:20000]
    assert len(idx) > 1000
    for i in idx[:5000]:
        t64, _, _ = oracle.intersect_fp64(s.tris[h["prim"][i]], rays[i])
        assert abs(h["t"][i] - t64) <= 1e-5 * abs(t64)
    # oracle traversal of the engine's own tree is bit-exact
    d = g.download()
    sub = rays[:50000]
    assert hits_equal(h[:50000], oracle.trace_final(d["nodes"], d["tri_ordered"], sub)).all()


def test_reference_rules_classification(oracle, ctx):
    """Compares the engine's hits with the faithful reference tracer (MT test, SBvh bin boxes,
    src/acceleration/sbvh.cc:60-145) and classifies every disagreement."""
    s = scenes.icosphere_room()
    g = gpu_bvh(ctx, s.tris)
    rays = scenes.camera_rays(s.camera, 128, 128)
    h = g.trace(rays)
    ref = oracle.SBvh(s.tris).trace(rays, mode=0)
    diff = np.nonzero(h["prim"] != ref["prim"])[0]
    unexplained = 0
    for i in diff:
        a, b = h[i], ref[i]
        if a["prim"] >= 0 and b["prim"] >= 0 and abs(a["t"] - b["t"]) <= 1e-6 * max(1.0, abs(a["t"])):
            continue  # tie
        if a["prim"] >= 0:
            b0 = 1 - a["b1"] - a["b2"]
            if min(a["b1"], a["b2"], b0) < 1e-6:
                continue  # grazing
        # reference misses: bin boxes not enclosing triangles, or the |det| < 1e-6 rejection
        fixed = oracle.SBvh(s.tris, fix_child_bounds=True).trace(rays[i:i + 1], mode=1)
        if fixed["prim"][0] == a["prim"]:
            continue
        unexplained += 1
    assert unexplained == 0
    assert len(diff) < 0.01 * len(rays)


def test_reference_brute_force_classification(oracle, ctx):
    """Compares the engine's hits with the reference's own brute force (MT with |det| < 1e-6
    rejected, strict closest hit: src/acceleration/dummy_acceleration_structure.cc:21-56) on
    the icosphere room and the mixed scene's small triangles, and classifies every
    disagreement as a tie, a grazing hit or a determinant rejection."""
    for name in ("icosphere", "mixed"):
        if name == "icosphere":
            s = scenes.icosphere_room()
            rays = scenes.camera_rays(s.camera, 128, 128)
        else:
            s = scenes.mixed_2m(n_small=20_000, n_thin=3_000)
            rays = scenes.random_rays_towards(s.tris, 8192, seed=21)
        say(f"reference brute force vs engine: {name}")
        h = gpu_bvh(ctx, s.tris).trace(rays)
        ref = oracle.trace_dummy(s.tris, rays)
        diff = np.nonzero(h["prim"] != ref["prim"])[0]
        kinds = {"tie": 0, "grazing": 0, "det": 0}
        for i in diff:
            a, b = h[i], ref[i]
            if a["prim"] >= 0 and b["prim"] >= 0 and abs(a["t"] - b["t"]) <= 1e-6 * max(1.0, abs(a["t"])):
                kinds["tie"] += 1
                continue
            graze = False
            for x, cond in ((a, a["prim"] >= 0), (b, b["prim"] >= 0)):
                if cond and min(x["b1"], x["b2"], 1 - x["b1"] - x["b2"]) < 1e-6:
                    graze = True
            if graze:
                kinds["grazing"] += 1
                continue
            # the engine's triangle is rejected by the reference's test: |det| < 1e-6
            assert a["prim"] >= 0, (name, i)
            tri = s.tris[a["prim"]].astype(np.float64).reshape(3, 3)
            d = rays[i, 4:7].astype(np.float64)
            det = np.dot(tri[1] - tri[0], np.cross(d, tri[2] - tri[0]))
            assert abs(det) < 1e-6 * (1 + 1e-5), (name, i, det)
            assert oracle.intersect_mt(s.tris[a["prim"]], rays[i])[0] == np.finfo(np.float32).max
            kinds["det"] += 1
        say(f"  {len(diff)} of {len(rays)} differ: {kinds}")
        assert len(diff) < 0.02 * len(rays)


@pytest.mark.parametrize("bv", [0, 1])
def test_trace_conservative_for_in_plane_and_grazing_rays(oracle, ctx, bv):
    """Rays that run in a wall's plane or graze a surface at their origin: the two-child, wide
    and counting traces all give the brute-force hits."""
    from diplodocus_amd import _lib

    from .test_oracle_sbvh import cornell_surface_rays

    tris, rays, shadow = cornell_surface_rays(60000, 3)
    say(f"in-plane rays bv={bv}")
    brute = oracle.trace_brute(tris, rays)
    bany = oracle.trace_brute(tris, shadow, 1)
    g = gpu_bvh(ctx, tris, bv=bv)
    assert hits_equal(g.trace(rays), brute).all()
    assert ((g.trace(shadow, 1)["prim"] >= 0) == (bany["prim"] >= 0)).all()
    h, _, _ = g.trace_counters(rays)
    assert hits_equal(h, brute).all()
    g.widen(4)
    assert hits_equal(g.trace(rays, _lib.DPL_TRACE_WIDE), brute).all()
    g.close()


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 9])
def test_trace_tiny_scenes(oracle, ctx, n):
    """Scenes whose whole tree is one leaf (the root's right slot an empty leaf) or a few
    nodes: every trace form equals brute force."""
    from diplodocus_amd import _lib

    rng = np.random.default_rng(n)
    tris = (rng.uniform(0, 1, (n, 1, 3)) + rng.normal(0, 0.3, (n, 3, 3))).reshape(n, 9).astype(np.float32)
    rays = scenes.random_rays_towards(tris, 4096, seed=n, jitter=0.4)
    say(f"tiny scene n={n}")
    for leaf in (1, 4, 15):
        g = gpu_bvh(ctx, tris, leaf=leaf)
        for flags in (0, 1, 2):
            brute = oracle.trace_brute(tris, rays, flags)
            h = g.trace(rays, flags)
            if flags & 1:
                assert ((h["prim"] >= 0) == (brute["prim"] >= 0)).all(), (leaf, flags)
            else:
                assert hits_equal(h, brute).all(), (leaf, flags)
        g.widen(4)
        assert hits_equal(g.trace(rays, _lib.DPL_TRACE_WIDE), oracle.trace_brute(tris, rays)).all(), leaf
        g.close()


def test_trace_error_paths(ctx):
    from diplodocus_amd import accel

    g = gpu_bvh(ctx, soup(10, 1))
    with pytest.raises(accel.DplError):
        g.trace(np.zeros((4, 8), np.float32), flags=16)
    assert len(g.trace(np.zeros((0, 8), np.float32))) == 0


def test_coherent_order_gives_the_same_hits(ctx):
    """DPL_TRACE_COHERENT sorts the rays before tracing: the hits at every ray's index equal
    the unsorted trace's bit for bit (closest, any-hit, culled, wide), including rays that miss
    the scene box, start inside it, are cut short by tmax or carry NaN, and small batches."""
    import torch
    from diplodocus_amd import _lib

    tris = scenes.mixed_2m(n_small=60_000, n_thin=8_000).tris
    g = gpu_bvh(ctx, tris, mt=1, bv=1)
    n = (1 << 20) + 4097
    rays = scenes.random_rays_towards(tris, n, seed=5)
    rng = np.random.default_rng(5)
    k = rng.choice(n, 3000, replace=False)
    rays[k[:1000], 4:7] *= -1.0  # aimed away from the scene: they miss the box
    v = tris.reshape(-1, 3)
    rays[k[1000:2000], 0:3] = v[rng.integers(0, len(v), 1000)] * 0.5 + v.mean(0) * 0.5  # inside the box
    rays[k[2000:2500], 7] = 0.5  # short spans
    rays[k[2500:2600], 4] = np.nan
    dev = torch.from_numpy(rays).cuda()
    for flags in (0, _lib.DPL_TRACE_ANY, _lib.DPL_TRACE_CULL_BACK):
        a = g.trace(dev, flags | _lib.DPL_TRACE_COHERENT)
        b = g.trace(dev, flags)
        assert torch.equal(a.view(torch.int32), b.view(torch.int32)), flags
    small = dev[:777]
    a = g.trace(small, _lib.DPL_TRACE_COHERENT)
    assert torch.equal(a.view(torch.int32), g.trace(small).view(torch.int32))
    host = g.trace(rays[:5000], _lib.DPL_TRACE_COHERENT)  # host buffers
    assert hits_equal(host, g.trace(rays[:5000])).all()
    g.widen(4)
    a = g.trace(dev, _lib.DPL_TRACE_WIDE | _lib.DPL_TRACE_COHERENT)
    b = g.trace(dev, _lib.DPL_TRACE_WIDE)
    assert torch.equal(a.view(torch.int32), b.view(torch.int32))
    g.close()


def test_handles_outliving_their_context():
    """A BVH and a trace graph still alive when their context is destroyed are detached:
    calls on them return an error and their destroy frees only their own memory."""
    import torch
    from diplodocus_amd import accel

    c = accel.Context(0)
    g = gpu_bvh(c, soup(500, 3))
    rays = torch.from_numpy(scenes.random_rays_towards(soup(500, 3), 256, seed=2)).cuda()
    hits = torch.empty((256, 4), dtype=torch.float32, device="cuda")
    gr = accel.TraceGraph(g, rays, hits)
    gr.launch()
    c.close()
    with pytest.raises(accel.DplError):
        g.build(soup(10, 1), accel.make_params())
    with pytest.raises(accel.DplError):
        gr.launch()
    gr.close()
    g.close()
