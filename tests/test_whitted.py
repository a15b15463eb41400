"""Whitted ray tracing (the reference's CpuRaytracer, src/renderer/cpu_raytracer.cc:50-288):
the oracle's restatement pinned by closed forms (emission and background pass through,
full shadow, depth 0 drops reflections), its own power function against FP64, and the
weighted depth-first evaluation against the reference's recursion as written; the GPU
kernel bit-exact against the oracle on the Cornell room (36 triangles, 256 x 256, depth 3)
and on the icosphere room with vertex normals, SOBB bounds and back-face culling."""
import numpy as np
import pytest

from diplodocus_amd import scenes


def quad(p0, p1, p2, p3):
    return [p0 + p1 + p2, p0 + p2 + p3]


def test_pow_pos_against_fp64(oracle):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(1e-6, 1.0, 3000), np.geomspace(1e-30, 1.0, 300), [1.0, 0.5, 0.999999]])
    ys = np.concatenate([rng.uniform(0.0, 256.0, 3000), rng.uniform(0.0, 4.0, 300), [7.0, 64.0, 1e4]])
    # FP32 evaluation of exp2(y log2 x): the rounding of log2 x is amplified by y, so the
    # relative error grows with |z| = |y log2 x|; bound it by 4e-7 (1 + |z|)
    worst = 0.0
    for x, y in zip(xs.astype(np.float32), ys.astype(np.float32)):
        ref = float(x) ** float(y)
        got = oracle.pow_pos(float(x), float(y))
        if ref < 1e-37:  # results below the normal range flush to zero
            assert got == 0.0 or abs(got - ref) < 1e-37
            continue
        z = abs(float(y) * np.log2(float(x)))
        worst = max(worst, abs(got - ref) / ref / (1.0 + z))
    assert worst < 4e-7, worst
    assert oracle.pow_pos(0.3, 0.0) == 1.0 and oracle.pow_pos(0.0, 0.0) == 1.0
    assert oracle.pow_pos(0.0, 5.0) == 0.0 and oracle.pow_pos(1.0, 123.0) == 1.0


def test_area_light_samples_are_hash_uniforms(oracle):
    # key = xor of seed, px, py, light, sample, dimension times odd constants, then HashU32
    # and the top 24 bits (cpu_raytracer.cc:28-39; util.h:24-37)
    def expect(seed, px, py, light, s, dim):
        m = 0xFFFFFFFF
        key = 0
        for v, k in ((seed, 0x9E3779B1), (px, 0x165667B1), (py, 0xD3A2646C), (light, 0x85EBCA6B), (s, 0xC2B2AE35),
                     (dim, 0x27D4EB2F)):
            key ^= (v * k) & m
        return (oracle.hash_u32(key) >> 8) / 16777216.0

    for args in [(42, 0, 0, 0, 0, 0), (42, 17, 3, 1, 7, 1), (42, 255, 255, 0, 3, 0), (7, 1, 2, 3, 4, 5)]:
        assert oracle.area_light_sample01(*args) == np.float32(expect(*args))
    vals = np.array([oracle.area_light_sample01(42, x, 5, 0, s, 0) for x in range(64) for s in range(16)])
    assert 0.0 <= vals.min() and vals.max() < 1.0 and abs(vals.mean() - 0.5) < 0.03


def emitter_scene():
    """An emissive triangle straight ahead of the camera (material 1) and nothing else."""
    tris = np.array(quad([-1, -1, 2], [1, -1, 2], [1, 1, 2], [-1, 1, 2]), np.float32)
    mats = np.stack([scenes.whitted_material((0.5, 0.5, 0.5)),
                     scenes.whitted_material((0, 0, 0), emission=(3.0, 2.0, 1.0))])
    cam = scenes.Camera(pos=(0.0, 0.0, 0.0), dir=(0.0, 0.0, 1.0), up=(0.0, 1.0, 0.0), fov=1.2, far=100.0)
    return tris, np.array([1, 1], np.int32), mats, cam


def test_emission_and_background_pass_through(oracle):
    tris, mat, mats, cam = emitter_scene()
    p = oracle.Ploc(tris, 0, 16, 4)
    img = oracle.whitted(p, mat, mats, 64, 48, 3, 4, cam, background=(0.25, 0.5, 0.75))
    center = img[24, 32]
    assert (center == np.float32([3.0, 2.0, 1.0])).all()
    # the quad spans tan(angle) <= 0.5 of the view: the outer pixels see the background
    assert (img[0, 0] == np.float32([0.25, 0.5, 0.75])).all()
    assert (img[47, 63] == np.float32([0.25, 0.5, 0.75])).all()


def shadow_scene(blocker: bool):
    """A floor (y = 0) under a small light quad at y = 2 facing down; optionally a large
    opaque blocker at y = 1 between them."""
    tris = quad([-5, 0, -5], [-5, 0, 5], [5, 0, 5], [5, 0, -5])  # normal +y
    mat = [0, 0]
    tris += quad([-0.2, 2, -0.2], [0.2, 2, -0.2], [0.2, 2, 0.2], [-0.2, 2, 0.2])  # normal -y
    mat += [1, 1]
    if blocker:
        tris += quad([-3, 1, -3], [3, 1, -3], [3, 1, 3], [-3, 1, 3])
        mat += [2, 2]
    mats = np.stack([scenes.whitted_material((0.8, 0.8, 0.8)),
                     scenes.whitted_material((0, 0, 0), emission=(10.0, 10.0, 10.0)),
                     scenes.whitted_material((0.5, 0.5, 0.5))])
    cam = scenes.Camera(pos=(0.0, 0.5, -1.0), dir=(0.0, -0.45, 1.0), up=(0.0, 1.0, 0.0), fov=0.6, far=100.0)
    return np.array(tris, np.float32), np.array(mat, np.int32), mats, cam


def test_full_shadow_and_direct_light(oracle):
    lit_t, lit_m, mats, cam = shadow_scene(False)
    img = oracle.whitted(oracle.Ploc(lit_t, 0, 16, 4), lit_m, mats, 32, 32, 0, 8, cam)
    assert (img > 0).all(), "the floor in view is lit"
    sh_t, sh_m, mats, cam = shadow_scene(True)
    # the camera is below the blocker: every floor point sees the light only through it
    img = oracle.whitted(oracle.Ploc(sh_t, 0, 16, 4), sh_m, mats, 32, 32, 0, 8, cam)
    assert (img == 0).all(), "every shadow ray is blocked"


def test_depth_zero_drops_secondary_rays(oracle):
    s, mats = scenes.cornell_whitted()
    p = oracle.Ploc(s.tris, 0, 16, 4)
    d0 = oracle.whitted(p, s.mat_id, mats, 64, 64, 0, 4, s.camera)
    d3 = oracle.whitted(p, s.mat_id, mats, 64, 64, 3, 4, s.camera)
    plain = mats[:, 3:9].sum(1) == 0  # no specular, no transmittance
    # pixels whose primary hit is a plain material have no secondary rays
    hits = oracle.trace_brute(s.tris, scenes.camera_rays(s.camera, 64, 64))
    prim = hits["prim"].reshape(64, 64)
    plain_px = (prim >= 0) & plain[s.mat_id[np.maximum(prim, 0)]]
    assert (d0[plain_px] == d3[plain_px]).all()
    assert (d3[~plain_px & (prim >= 0)] >= d0[~plain_px & (prim >= 0)] - 1e-6).all()
    assert (d3 != d0).any()


@pytest.mark.parametrize("cull", [False, True])
def test_weighted_sum_matches_the_reference_recursion(oracle, cull):
    s, mats = scenes.cornell_whitted()
    p = oracle.Ploc(s.tris, 1, 16, 4)
    a = oracle.whitted(p, s.mat_id, mats, 128, 128, 4, 8, s.camera, cull=cull)
    b = oracle.whitted(p, s.mat_id, mats, 128, 128, 4, 8, s.camera, cull=cull, recursive=True)
    np.testing.assert_allclose(a, b, rtol=2e-6, atol=1e-7)


def test_ray_counters_closed_forms(oracle):
    """The reference's counters (cpu_raytracer.cc:103,131,147,223): one primary ray per pixel;
    at depth 0 no secondary rays and, per pixel whose primary hit is not emissive, one shadow
    ray per light and sample; every trace is a query; the recursion as written and the weighted
    depth-first evaluation count the same."""
    s, mats = scenes.cornell_whitted()
    p = oracle.Ploc(s.tris, 0, 16, 4)
    W = H = 48
    _, c0 = oracle.whitted(p, s.mat_id, mats, W, H, 0, 3, s.camera, counts=True)
    hits = oracle.trace_brute(s.tris, scenes.camera_rays(s.camera, W, H))["prim"]
    emissive = mats[:, 9:12].sum(1) > 0
    lit = int(((hits >= 0) & ~emissive[s.mat_id[np.maximum(hits, 0)]]).sum())
    assert c0["primary_ray_count"] == W * H and c0["secondary_ray_count"] == 0
    assert c0["shadow_ray_count"] == lit * 2 * 3
    # no light sample lies within 1e-6 of a shaded point here, so every shadow ray is traced
    assert c0["query_count"] == W * H + c0["shadow_ray_count"]
    _, c4 = oracle.whitted(p, s.mat_id, mats, W, H, 4, 3, s.camera, counts=True)
    _, r4 = oracle.whitted(p, s.mat_id, mats, W, H, 4, 3, s.camera, counts=True, recursive=True)
    assert c4 == r4 and c4["secondary_ray_count"] > 0
    assert c4["query_count"] == c4["primary_ray_count"] + c4["secondary_ray_count"] + c4["shadow_ray_count"]
    assert c4["traversal_count"] > c0["traversal_count"] and c4["intersection_count"] > c0["intersection_count"]


def icosphere_whitted():
    s = scenes.icosphere_room()
    mats = np.stack([scenes.whitted_material(tuple(s.albedo[m]),
                                             specular=(0.3, 0.3, 0.3) if m in (7, 9, 12) else (0, 0, 0),
                                             transmittance=(0.7, 0.7, 0.7) if m in (8, 13) else (0, 0, 0),
                                             emission=tuple(s.emission[m]), shininess=32.0,
                                             ior=1.33 if m in (8, 13) else 1.0) for m in range(len(s.albedo))])
    vn = scenes.sphere_vertex_normals(s.tris, 36, (0.68, 0.6, 0.45))
    return s, mats, vn


def _gpu(arr):
    import torch

    return torch.from_numpy(np.ascontiguousarray(arr)).cuda()


@pytest.mark.gpu
def test_gpu_whitted_cornell_bitexact(oracle, ctx):
    from diplodocus_amd import accel, render

    s, mats = scenes.cornell_whitted()
    g = accel.GpuBvh(ctx).build(s.tris, accel.make_params(0, 16, 4))
    lights = render.emissive_triangles(s.mat_id, mats)
    assert len(lights) == 2
    wp = render.make_whitted_params(256, 256, 3, 8, s.camera)
    img = render.render_whitted(g, wp, _gpu(s.tris), _gpu(s.mat_id), _gpu(mats), _gpu(lights)).cpu().numpy()
    ref = oracle.whitted(oracle.Ploc(s.tris, 0, 16, 4), s.mat_id, mats, 256, 256, 3, 8, s.camera)
    assert (img.view(np.uint32) == ref.view(np.uint32)).all(), np.abs(img - ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("scene,bv", [("cornell", 0), ("icosphere", 1)])
def test_gpu_whitted_stats_match_oracle(oracle, ctx, scene, bv):
    """dpl_render_whitted_stats: the image is the same as without counters, and the ray
    counters and traversal work equal the oracle's counts exactly."""
    from diplodocus_amd import accel, render

    if scene == "cornell":
        s, mats = scenes.cornell_whitted()
        vn = None
    else:
        s, mats, vn = icosphere_whitted()
    g = accel.GpuBvh(ctx).build(s.tris, accel.make_params(morton_type=1, radius=16, max_leaf=4, bv_type=bv))
    lights = render.emissive_triangles(s.mat_id, mats)
    wp = render.make_whitted_params(96, 80, 4, 4, s.camera)
    args = (_gpu(s.tris), _gpu(s.mat_id), _gpu(mats), _gpu(lights))
    plain = render.render_whitted(g, wp, *args, vertex_normals=None if vn is None else _gpu(vn)).cpu().numpy()
    img, st = render.render_whitted(g, wp, *args, vertex_normals=None if vn is None else _gpu(vn), stats=True)
    img = img.cpu().numpy()
    assert (img.view(np.uint32) == plain.view(np.uint32)).all()
    # the same tree as the GPU's (bit-exact build), so the traversal work is comparable
    ref, rc = oracle.whitted(oracle.Ploc(s.tris, 1, 16, 4, bv_type=bv), s.mat_id, mats, 96, 80, 4, 4, s.camera,
                             vertex_normals=vn, counts=True)
    assert (img.view(np.uint32) == ref.view(np.uint32)).all()
    got = {k: st[k] for k in rc}
    assert got == rc, (got, rc)
    assert st["raytracing_ms"] > 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("bv,cull", [(0, False), (1, True)])
def test_gpu_whitted_icosphere_vertex_normals(oracle, ctx, bv, cull):
    from diplodocus_amd import accel, render

    s, mats, vn = icosphere_whitted()
    g = accel.GpuBvh(ctx).build(s.tris, accel.make_params(morton_type=1, radius=16, max_leaf=4, bv_type=bv))
    lights = render.emissive_triangles(s.mat_id, mats)
    wp = render.make_whitted_params(160, 120, 4, 4, s.camera, background=(0.1, 0.2, 0.3), backface_culling=cull)
    img = render.render_whitted(g, wp, _gpu(s.tris), _gpu(s.mat_id), _gpu(mats), _gpu(lights),
                                vertex_normals=_gpu(vn)).cpu().numpy()
    ref = oracle.whitted(oracle.Ploc(s.tris, 1, 16, 4), s.mat_id, mats, 160, 120, 4, 4, s.camera,
                         background=(0.1, 0.2, 0.3), cull=cull, vertex_normals=vn)
    assert (img.view(np.uint32) == ref.view(np.uint32)).all(), np.abs(img - ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("side", [384, 512, 768])
def test_gpu_whitted_lanes_per_pixel(oracle, ctx, side):
    """Larger images run with 4, 2 and 1 lanes per pixel on a 256-CU GPU (small ones with 8):
    image and counters equal the oracle's for a sample count that is no multiple of the
    lanes (3 samples per light)."""
    from diplodocus_amd import accel, render

    s, mats = scenes.cornell_whitted()
    g = accel.GpuBvh(ctx).build(s.tris, accel.make_params(0, 16, 4))
    lights = render.emissive_triangles(s.mat_id, mats)
    wp = render.make_whitted_params(side, side, 3, 3, s.camera)
    img, st = render.render_whitted(g, wp, _gpu(s.tris), _gpu(s.mat_id), _gpu(mats), _gpu(lights), stats=True)
    img = img.cpu().numpy()
    ref, rc = oracle.whitted(oracle.Ploc(s.tris, 0, 16, 4), s.mat_id, mats, side, side, 3, 3, s.camera, counts=True)
    assert (img.view(np.uint32) == ref.view(np.uint32)).all(), np.abs(img - ref).max()
    assert {k: st[k] for k in rc} == rc
