"""Seeded workload generators: exact counts, golden hashes, closed-mesh property."""
import json
import os

import numpy as np

from diplodocus_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))


def golden():
    return json.load(open(os.path.join(HERE, "golden", "scenes.json")))


def test_golden_hashes():
    from tests.golden.make_scene_hashes import fixtures

    assert fixtures() == golden()


def test_counts():
    assert scenes.cornell_room().n == 36
    s = scenes.icosphere_room()
    assert s.n == 36 + 69_632
    assert s.albedo.shape == (16, 3) and (s.albedo >= 0.1).all() and (s.albedo <= 0.9).all()
    assert (s.emission.sum(1) > 0).sum() == 1
    assert set(np.unique(s.mat_id)) == set(range(16))


def test_icosphere_is_closed():
    tris = scenes.displaced_icosphere(59, 6, (0, 0, 0), 1.0, 0.15, 7)
    assert tris.shape[0] == 69_632
    v = tris.reshape(-1, 3)
    _, idx = np.unique(v, axis=0, return_inverse=True)
    f = idx.reshape(-1, 3)
    e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert (cnt == 2).all()  # every edge shared by exactly two triangles
    nv = idx.max() + 1
    assert nv - len(cnt) + len(f) == 2  # Euler characteristic of a sphere


def test_camera_rays_follow_reference_model():
    cam = scenes.Camera()
    r = scenes.camera_rays(cam, 8, 6)
    d = r[:, 4:7].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    # the centre of the image plane looks along the camera direction (no pixel offset)
    mid = (r[0, 4:7] + r[-1, 4:7]) / 2
    assert mid[0] == 0 and abs(mid[1]) < 1e-6
    assert (r[:, 3] == np.float32(1e-6)).all() and (r[:, 7] == np.float32(cam.far)).all()
