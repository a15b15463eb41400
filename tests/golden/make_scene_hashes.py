"""Writes tests/golden/scenes.json: triangle counts and SHA-256 prefixes of the seeded
workload buffers (triangles, material ids, rays), so that a generator change is caught.

Run: python tests/golden/make_scene_hashes.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from diplodocus_amd import scenes  # noqa: E402


def fixtures() -> dict:
    out = {}
    c1 = scenes.cornell_room()
    out["cornell_room"] = {"n": c1.n, "tris": scenes.buffer_hash(c1.tris), "mat": scenes.buffer_hash(c1.mat_id),
                           "camera_rays_32": scenes.buffer_hash(scenes.camera_rays(c1.camera, 32, 32))}
    c2 = scenes.icosphere_room()
    out["icosphere_room"] = {"n": c2.n, "tris": scenes.buffer_hash(c2.tris), "mat": scenes.buffer_hash(c2.mat_id),
                             "albedo": scenes.buffer_hash(c2.albedo)}
    c3 = scenes.mixed_2m()
    out["mixed_2m"] = {"n": c3.n, "tris": scenes.buffer_hash(c3.tris),
                       "rays_4096": scenes.buffer_hash(scenes.random_rays_towards(c3.tris, 4096, seed=11))}
    return out


if __name__ == "__main__":
    path = os.path.join(HERE, "scenes.json")
    with open(path, "w") as f:
        json.dump(fixtures(), f, indent=1, sort_keys=True)
    print("wrote", path)
