"""Scene loading (the reference's ObjSceneLoader, src/io/scene/obj_scene_loader.cc:20-204,
over its vendored tinyobjloader): the product's C++ loader (libdpl_scene.so through
diplodocus_amd.scene_io) and the oracle's restatement (oracle/objload.py) against

  * golden vectors written by the reference's own parser (tests/golden/make_obj_golden.py:
    tinyobjloader compiled from the reference's sources, oracle/_ref) for the reference's
    Cornell box scenes and a hand-written file of edge cases, bit for bit;
  * that parser live, on seeded random scenes, where oracle/_ref is built;
  * number spellings whose values the parsing rules fix, closed forms of the polygon split
    (a convex n-gon gives n - 2 triangles covering its area; a quad splits on its shorter
    diagonal), and the failures the loader reports;

then the loaded Cornell boxes are built and rendered on the GPU (Whitted, bit-exact against
the oracle's renderer on the same arrays)."""
import math
import os
import shutil

import numpy as np
import pytest

from oracle import objload
from tests import objgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OBJ_DIR = os.path.join(ROOT, "tests", "golden", "obj")
REF_DUMP = os.path.join(ROOT, "oracle", "_ref", "ref_obj_dump")
SCENES = ["CornellBox-Original", "CornellBox-Sphere", "edge_cases"]


def scene_io():
    from diplodocus_amd import scene_io as s

    return s


def bits_equal(a, b_u32) -> bool:
    a = np.ascontiguousarray(a, np.float32)
    return a.shape == b_u32.shape and (a.view(np.uint32) == b_u32).all()


def assert_matches_golden(s, g, normals, camera):
    assert bits_equal(s.tris, g["tris"])
    assert bits_equal(normals, g["normals"])
    assert (s.mat_id == g["mat_id"]).all()
    assert bits_equal(s.materials, g["materials"])
    assert list(s.material_names) == list(g["material_names"])
    assert (s.light_tri == g["light_tri"]).all()
    assert bits_equal(s.light_area, g["light_area"])
    assert bits_equal(camera, g["camera"])
    assert bits_equal(s.point_lights, g["point_lights"])


@pytest.mark.parametrize("name", SCENES)
def test_oracle_matches_reference_golden(name):
    g = np.load(os.path.join(OBJ_DIR, name + ".golden.npz"))
    s = objload.load_scene(OBJ_DIR, name)
    assert (s.has_normals == g["has_normals"].astype(bool)).all()
    assert_matches_golden(s, g, s.normals, s.camera)


@pytest.mark.parametrize("name", SCENES)
def test_loader_matches_reference_golden(name):
    g = np.load(os.path.join(OBJ_DIR, name + ".golden.npz"))
    s = scene_io().load_scene(OBJ_DIR, name)
    assert (s.has_vertex_normals == g["has_normals"].astype(bool)).all()
    assert_matches_golden(s, g, s.vertex_normals, s.camera_array)
    # the render rows are the same materials in dpl_material order
    m = s.materials
    assert (s.render_materials[:, 0:12] == m[:, 3:15]).all() and (s.render_materials[:, 12:15] == m[:, 15:18]).all()
    # area light colours are the emitting material's emission
    assert (s.light_color == m[s.mat_id[s.light_tri], 12:15]).all()


def test_cornell_box_contents():
    s = scene_io().load_scene(OBJ_DIR, "CornellBox-Original")
    assert s.n == 34 and len(s.material_names) == 8 and len(s.light_tri) == 2
    assert s.material_names[s.mat_id[s.light_tri[0]]] == "light"
    # the light is the 130 x 105 quad at y = 548.7 (CornellBox-Original.obj): two triangles of
    # half its area, emission (17, 12, 4)
    assert np.allclose(s.light_area, 0.5 * 130 * 105, rtol=1e-6)
    assert (s.light_color == np.float32([17, 12, 4])).all()
    cam = s.camera
    assert cam.pos == (278.0, 273.0, -1000.0) and cam.dir == (0.0, 0.0, 1.0) and cam.up == (0.0, 1.0, 0.0)
    assert s.point_lights.shape == (1, 7) and s.point_lights[0, 6] == 100000.0
    # geometric normals: the unit normal of each triangle's plane (Normalize(Cross(e1, e2)))
    t = s.tris.reshape(-1, 3, 3).astype(np.float64)
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    assert np.allclose(s.geom_normal, n, atol=1e-6)
    # r_ior = 1 / ior
    assert (s.materials[:, 17] == np.float32(1) / s.materials[:, 16]).all()


@pytest.mark.parametrize("seed", range(40))
def test_loader_matches_oracle_random(tmp_path, seed):
    objgen.write_scene(str(tmp_path), "r", seed)
    o = objload.load_scene(str(tmp_path), "r")
    p = scene_io().load_scene(str(tmp_path), "r")
    assert p.tris.view(np.uint32).tolist() == o.tris.view(np.uint32).tolist()
    assert (p.vertex_normals.view(np.uint32) == o.normals.view(np.uint32)).all()
    assert (p.has_vertex_normals == o.has_normals).all() and (p.mat_id == o.mat_id).all()
    assert (p.materials.view(np.uint32) == o.materials.view(np.uint32)).all()
    assert (p.light_tri == o.light_tri).all() and (p.light_area.view(np.uint32) == o.light_area.view(np.uint32)).all()
    assert (p.camera_array.view(np.uint32) == o.camera.view(np.uint32)).all()
    assert (p.point_lights.view(np.uint32) == o.point_lights.view(np.uint32)).all()


@pytest.mark.skipif(not os.path.exists(REF_DUMP), reason="oracle/_ref is built only where the reference is present")
@pytest.mark.parametrize("seed", range(1000, 1030))
def test_oracle_matches_reference_parser_live(tmp_path, seed):
    import sys

    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_obj_golden as G

    objgen.write_scene(str(tmp_path), "r", seed)
    g = G.to_arrays(G.dump([str(tmp_path), "r"]))
    assert g["ok"]
    s = objload.load_scene(str(tmp_path), "r")
    assert_matches_golden(s, g, s.normals, s.camera)


def write_vertex_file(path, spelling):
    with open(path, "w") as f:
        f.write(f"v {spelling} 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")


# value of the first coordinate for each spelling: the longest decimal prefix rounded to
# double then float; nan -> 0; inf -> largest double -> float infinity
SPELLINGS = [("0x10", 0.0), (".5", 0.5), ("5.", 5.0), ("+3", 3.0), ("-nan", 0.0), ("inf", np.inf),
             ("-Infinity", -np.inf), ("1e", 1.0), ("1e+", 1.0), ("1.5abc", 1.5), ("1e-45", np.float32(1e-45)),
             ("3.4028236e38", np.inf), ("00012", 12.0), ("1E2", 100.0), ("1.0d3", 1.0), ("1,5", 1.0),
             ("1e-320", 0.0), ("-0", -0.0), ("0.1", np.float32(0.1)),
             # double rounding: a hair above 1 + 2^-24 (the float tie) rounds to the tie in double,
             # then to even in float
             ("1.00000005960464477539930763", np.float32(1.0))]
BAD = ["1e-400", "1e400", "2e-324", "1.7976931348623159e308", "-.e1", ".", "abc", "+-1"]


@pytest.mark.parametrize("spelling,value", SPELLINGS)
def test_number_spellings(tmp_path, spelling, value):
    p = str(tmp_path / "n.obj")
    write_vertex_file(p, spelling)
    want = np.float32(value)
    for got in (objload.load_obj_scene(p).tris[0, 0], scene_io().load_obj(p).tris[0, 0]):
        assert np.float32(got).view(np.uint32) == want.view(np.uint32), (spelling, got, want)


@pytest.mark.parametrize("spelling", BAD)
def test_bad_numbers_fail(tmp_path, spelling):
    p = str(tmp_path / "n.obj")
    write_vertex_file(p, spelling)
    with pytest.raises(objload.ObjError):
        objload.load_obj_scene(p)
    with pytest.raises(scene_io().SceneLoadError, match="malformed vertex"):
        scene_io().load_obj(p)


def polygon_file(path, pts, extra=""):
    with open(path, "w") as f:
        for p in pts:
            f.write("v %r %r %r\n" % tuple(float(np.float32(c)) for c in p))
        f.write("f " + " ".join(str(i + 1) for i in range(len(pts))) + "\n" + extra)


@pytest.mark.parametrize("n", [5, 6, 7, 9, 12, 16])
def test_convex_polygon_split_covers_it(tmp_path, n):
    """A convex polygon splits into n - 2 triangles of its corners whose areas sum to the
    polygon's (shoelace) area."""
    rng = np.random.default_rng(n)
    ang = np.sort(rng.uniform(0, 2 * np.pi, n))
    pts = np.stack([3 * np.cos(ang), np.full(n, 1.5), 2 * np.sin(ang)], 1).astype(np.float32)
    p = str(tmp_path / "poly.obj")
    polygon_file(p, pts)
    x, z = pts[:, 0].astype(np.float64), pts[:, 2].astype(np.float64)
    shoelace = 0.5 * abs(np.dot(x, np.roll(z, -1)) - np.dot(z, np.roll(x, -1)))
    for s in (objload.load_obj_scene(p), scene_io().load_obj(p)):
        assert len(s.tris) == n - 2
        t = s.tris.reshape(-1, 3, 3).astype(np.float64)
        area = 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1).sum()
        assert math.isclose(area, shoelace, rel_tol=1e-5)
        corners = {tuple(c) for c in pts.tolist()}
        assert all(tuple(v) in corners for v in s.tris.reshape(-1, 3).tolist())


@pytest.mark.parametrize("short02", [True, False])
def test_quad_splits_on_the_shorter_diagonal(tmp_path, short02):
    # a kite whose 0-2 diagonal is shorter (or longer) than 1-3
    pts = np.array([[0, 0, 0], [1, -2, 0], [0, -1 if short02 else -4, 0], [-1, -2, 0]], np.float32)
    p = str(tmp_path / "q.obj")
    polygon_file(p, pts)
    want = [(0, 1, 2), (0, 2, 3)] if short02 else [(0, 1, 3), (1, 2, 3)]
    for s in (objload.load_obj_scene(p), scene_io().load_obj(p)):
        got = [tuple(int(np.where((pts == v).all(1))[0][0]) for v in t.reshape(3, 3)) for t in s.tris]
        assert got == want


def test_failures_are_reported(tmp_path):
    sio = scene_io()
    cases = {
        "zero.obj": ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n", "invalid face index"),
        "ahead.obj": ("v 0 0 0\nv 1 0 0\nf 1 2 3\nv 0 1 0\n", "beyond"),
        "rel.obj": ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf -4 -2 -1\n", "invalid face index"),
        "vn.obj": ("v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//2 3//1\n", "beyond"),
        "badmtl.obj": ("mtllib bad.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n", "material library"),
    }
    (tmp_path / "bad.mtl").write_text("newmtl a\nKd 0.5 x 0.5\n")
    for fname, (text, msg) in cases.items():
        p = str(tmp_path / fname)
        with open(p, "w") as f:
            f.write(text)
        with pytest.raises(objload.ObjError):
            objload.load_obj_scene(p)
        with pytest.raises(sio.SceneLoadError, match=msg):
            sio.load_obj(p)
    with pytest.raises(sio.SceneLoadError, match="cannot read"):
        sio.load_scene(str(tmp_path), "missing")
    shutil.copy(os.path.join(OBJ_DIR, "CornellBox-Original.obj"), tmp_path / "c.obj")
    shutil.copy(os.path.join(OBJ_DIR, "CornellBox-Original.mtl"), tmp_path / "CornellBox-Original.mtl")
    (tmp_path / "c.meta").write_text("cam_pos 0 1 2\ncam_zoom 3\n")
    with pytest.raises(sio.SceneLoadError, match="unrecognized metadata token 'cam_zoom'"):
        sio.load_scene(str(tmp_path), "c")
    with pytest.raises(objload.ObjError):
        objload.load_scene(str(tmp_path), "c")
    (tmp_path / "c.meta").write_text("cam_fov\n")
    with pytest.raises(sio.SceneLoadError, match="needs numbers"):
        sio.load_scene(str(tmp_path), "c")


def test_no_material_and_unknown_material(tmp_path):
    p = str(tmp_path / "m.obj")
    with open(p, "w") as f:
        f.write("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\nusemtl nothing\nf 1 2 3\n")
    for s in (objload.load_obj_scene(p), scene_io().load_obj(p)):
        assert list(s.mat_id) == [-1, -1] and len(s.light_tri) == 0


def _gpu(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("name,bv,depth", [("CornellBox-Original", 0, 8), ("CornellBox-Sphere", 1, 8),
                                           ("CornellBox-Sphere", 0, 3)])
def test_gpu_whitted_of_loaded_scene_bitexact(oracle, ctx, name, bv, depth):
    """The reference's default scene path: load, build, Whitted render (its sample config:
    max depth 8, 8 area-light samples) - GPU against the oracle on the loaded arrays."""
    from diplodocus_amd import accel, render

    s = scene_io().load_scene(OBJ_DIR, name)
    assert (s.mat_id >= 0).all()
    g = accel.GpuBvh(ctx).build(s.tris, accel.make_params(morton_type=1, radius=16, max_leaf=4, bv_type=bv))
    lights = render.emissive_triangles(s.mat_id, s.render_materials)
    assert (lights == s.light_tri).all()
    wp = render.make_whitted_params(128, 128, depth, 8, s.camera)
    vn = s.vertex_normals_or_none()
    img = render.render_whitted(g, wp, _gpu(s.tris), _gpu(s.mat_id), _gpu(s.render_materials), _gpu(lights),
                                vertex_normals=None if vn is None else _gpu(vn)).cpu().numpy()
    ref = oracle.whitted(oracle.Ploc(s.tris, 1, 16, 4), s.mat_id, s.render_materials, 128, 128, depth, 8, s.camera,
                         vertex_normals=vn)
    assert (img.view(np.uint32) == ref.view(np.uint32)).all(), np.abs(img - ref).max()
    assert img.sum() > 0
