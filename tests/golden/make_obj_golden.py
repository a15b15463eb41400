"""Writes the scene-loading golden vectors tests/golden/obj/<name>.golden.npz from the
reference's own OBJ parser (oracle/_ref/ref_obj_dump: the vendored tinyobjloader compiled
from the reference's sources by `make -C oracle ref`, with the reference's LoadObj /
LoadMetadata conversion restated around it).

Every float is kept as its IEEE-754 bit pattern (uint32), so the fixtures pin bit-exact
results. Run from the repository root where the reference is present:
    make -C oracle ref && python tests/golden/make_obj_golden.py
"""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DUMP = os.path.join(ROOT, "oracle", "_ref", "ref_obj_dump")
SCENES = ["CornellBox-Original", "CornellBox-Sphere", "edge_cases"]


def dump(args) -> dict:
    out = subprocess.run([DUMP] + list(args), check=True, capture_output=True, text=True).stdout
    return json.loads(out)


def to_arrays(d: dict) -> dict:
    tv = np.array([t["v"] for t in d["triangles"]], np.uint32).reshape(-1, 18)
    return {
        "ok": np.array(d["ok"]),
        "tris": tv[:, :9],
        "normals": tv[:, 9:],
        "has_normals": np.array([t["has_n"] for t in d["triangles"]], np.uint8),
        "mat_id": np.array([t["mat"] for t in d["triangles"]], np.int32),
        "materials": np.array([m["v"] for m in d["materials"]], np.uint32).reshape(-1, 19),
        "material_names": np.array([m["name"] for m in d["materials"]], dtype="U64"),
        "light_tri": np.array([a[0] for a in d["area_lights"]], np.int32),
        "light_area": np.array([a[1] for a in d["area_lights"]], np.uint32),
        "camera": np.array(d["camera"], np.uint32),
        "point_lights": np.array(d["point_lights"], np.uint32).reshape(-1, 7),
    }


def main() -> int:
    if not os.path.exists(DUMP):
        print(f"{DUMP} is missing: run `make -C oracle ref` where the reference is present", file=sys.stderr)
        return 1
    obj_dir = os.path.join(HERE, "obj")
    for name in SCENES:
        arrays = to_arrays(dump([obj_dir, name]))
        assert arrays["ok"], name
        path = os.path.join(obj_dir, name + ".golden.npz")
        np.savez_compressed(path, **arrays)
        print(f"{path}: {len(arrays['tris'])} triangles, {len(arrays['materials'])} materials, "
              f"{len(arrays['light_tri'])} area lights")
    return 0


if __name__ == "__main__":
    sys.exit(main())
