"""Seeded random OBJ/MTL/.meta text for the scene-loading tests: polygons of 3 to 12 corners
(convex, star-shaped, collinear runs, repeated corners) in random planes and windings, every
face index form with positive and negative indices, numbers in many spellings, materials
switched, unknown and absent, groups, objects, smoothing groups, comments, tabs and CRLF."""
from __future__ import annotations

import math
import os

import numpy as np


def _num(rng, x: float) -> str:
    k = rng.integers(0, 9)
    if k == 0:
        return f"{x:.6f}"
    if k == 1:
        return f"{x:.9g}"
    if k == 2:
        return f"{x:e}"
    if k == 3:
        return f"{x:+.4f}"
    if k == 4:  # ".5" / "-.5" spelling
        s = f"{x:.5f}"
        return s.replace("0.", ".", 1) if s.startswith("0.") or s.startswith("-0.") else s
    if k == 5:
        return f"{x:.3f}E0"
    if k == 6:
        return repr(float(np.float32(x)))
    if k == 7:
        return f"{x:.17g}"
    return f"{round(x, 2)}"


def _polygon(rng, n: int) -> np.ndarray:
    """n corners in the plane z = 0 before a random rotation: convex, star-shaped or with
    collinear / repeated corners."""
    kind = rng.integers(0, 4)
    ang = np.sort(rng.uniform(0, 2 * math.pi, n)) if kind != 0 else np.linspace(0, 2 * math.pi, n, endpoint=False)
    if kind == 1:  # star: alternating radii
        r = np.where(np.arange(n) % 2 == 0, 1.0, rng.uniform(0.2, 0.6))
    elif kind == 2:  # star-shaped with random radii
        r = rng.uniform(0.2, 1.5, n)
    else:
        r = np.ones(n)
    p = np.stack([r * np.cos(ang), r * np.sin(ang), np.zeros(n)], 1)
    if kind == 3 and n >= 5:  # a collinear run and a repeated corner
        p[2] = 0.5 * (p[1] + p[3])
        p[4] = p[3]
    if rng.random() < 0.5:
        p = p[::-1]
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if rng.random() < 0.3:  # axis-aligned planes too
        q = np.eye(3)[rng.permutation(3)]
    return p @ q.T * rng.uniform(0.5, 20.0) + rng.uniform(-50, 50, 3)


def write_scene(dirpath: str, name: str, seed: int, faces: int = 40) -> None:
    """Writes <dirpath>/<name>.obj, <name>.mtl and <name>.meta."""
    rng = np.random.default_rng(seed)
    nl = "\r\n" if rng.random() < 0.3 else "\n"
    mats = [f"m{i}" for i in range(int(rng.integers(1, 5)))]
    mtl = []
    for i, m in enumerate(mats):
        mtl.append(f"newmtl {m}")
        mtl.append("Kd " + " ".join(_num(rng, v) for v in rng.uniform(0, 1, 3)))
        if rng.random() < 0.5:
            mtl.append("\tKs " + " ".join(_num(rng, v) for v in rng.uniform(0, 1, 3)))
        if rng.random() < 0.3:
            mtl.append("Ke " + " ".join(_num(rng, v) for v in rng.uniform(0, 20, 3)))
        if rng.random() < 0.3:
            mtl.append(f"Ni {_num(rng, rng.uniform(1.0, 2.5))}")
        if rng.random() < 0.3:
            mtl.append(f"Tf {_num(rng, 0.9)} 0.8 0.7")
        if rng.random() < 0.3:
            mtl.append(f"Ns {_num(rng, rng.uniform(1, 500))}")
        if rng.random() < 0.2:
            mtl.append(f"Tr {_num(rng, rng.uniform(0, 1))}")
        if rng.random() < 0.2:
            mtl.append(f"d {_num(rng, rng.uniform(0, 1))}")
        if rng.random() < 0.2:
            mtl.append("map_Kd tex.png")
        mtl.append("")
    with open(os.path.join(dirpath, name + ".mtl"), "w", newline="") as f:
        f.write(nl.join(mtl) + nl)
    out = [f"# random scene {seed}", f"mtllib {name}.mtl"]
    nv = nn = nt = 0
    for fi in range(faces):
        n = int(rng.choice([3, 3, 4, 4, 5, 6, 7, 8, 10, 12]))
        pts = _polygon(rng, n)
        for p in pts:
            out.append("v " + " ".join(_num(rng, float(c)) for c in p))
        nv += n
        has_vn = rng.random() < 0.5
        has_vt = rng.random() < 0.4
        if has_vn:
            for _ in range(n):
                d = rng.normal(size=3)
                d /= np.linalg.norm(d)
                out.append("vn " + " ".join(_num(rng, float(c)) for c in d))
            nn += n
        if has_vt:
            for _ in range(n):
                out.append(f"vt {_num(rng, rng.random())} {_num(rng, rng.random())}")
            nt += n
        r = rng.random()
        if r < 0.15:
            out.append(f"usemtl {rng.choice(mats)}")
        elif r < 0.2:
            out.append("usemtl not_defined")
        elif r < 0.25:
            out.append(f"g group{fi}")
        elif r < 0.28:
            out.append(f"o object {fi}")
        elif r < 0.31:
            out.append(f"s {int(rng.integers(0, 3))}")
        elif r < 0.34:
            out.append("# a comment line")
        elif r < 0.36:
            out.append("")
        neg = rng.random() < 0.4
        corners = []
        for k in range(n):
            vi = (k - n) if neg else (nv - n + k + 1)
            s = str(vi)
            ti = (k - n) if neg else (nt - n + k + 1)
            ni = (k - n) if neg else (nn - n + k + 1)
            if has_vt and has_vn:
                s += f"/{ti}/{ni}"
            elif has_vn:
                s += f"//{ni}" if rng.random() < 0.9 else "//0"
            elif has_vt:
                s += f"/{ti}"
            corners.append(s)
        sep = "\t" if rng.random() < 0.1 else " "
        tail = "  # trailing comment" if rng.random() < 0.1 else ("  " if rng.random() < 0.1 else "")
        out.append("f " + sep.join(corners) + tail)
    with open(os.path.join(dirpath, name + ".obj"), "w", newline="") as f:
        f.write(nl.join(out) + nl)
    cam = rng.uniform(-5, 5, 9)
    meta = [f"cam_pos {_meta_num(rng, cam[0])} {_meta_num(rng, cam[1])} {_meta_num(rng, cam[2])}",
            f"cam_dir {_meta_num(rng, cam[3])} {_meta_num(rng, cam[4])} {_meta_num(rng, cam[5])}",
            f"cam_up 0 1 0", f"cam_fov {_meta_num(rng, rng.uniform(0.2, 1.5))}",
            f"cam_far {_meta_num(rng, rng.uniform(10, 5000))}"]
    for _ in range(int(rng.integers(0, 3))):
        meta.append("point_light " + " ".join(_meta_num(rng, v) for v in rng.uniform(-10, 10, 7)))
    order = rng.permutation(len(meta))
    with open(os.path.join(dirpath, name + ".meta"), "w") as f:
        f.write("\n".join(meta[i] for i in order) + "\n")


def _meta_num(rng, x: float) -> str:
    k = rng.integers(0, 4)
    if k == 0:
        return f"{x:.7g}"
    if k == 1:
        return f"{x:.12g}"
    if k == 2:
        return f"{x:e}"
    return f"{x:.3f}"
