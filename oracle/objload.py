"""Oracle (test infrastructure only; nothing in the product imports it): a restatement of the
reference's scene loading in plain Python, used to check the product's C++ loader.

The reference loads a scene as <dir>/<name>.meta then <dir>/<name>.obj
(ObjSceneLoader::Load, src/io/scene/obj_scene_loader.cc:188-204). The OBJ/MTL parsing is done
by its vendored tinyobjloader (third-party/tinyobjloader/tiny_obj_loader.h, version 2.x with
the fast_float number parser, used with the default ObjReaderConfig: triangulate = true,
triangulation_method "simple"); LoadObj (obj_scene_loader.cc:20-126) then turns tinyobj's
shapes into triangles, materials and area lights, and LoadMetadata (:128-186) reads the
camera and point lights. Restated here:

  * numbers: tiny_obj_loader.h tryParseDouble (fast_float path, :5964-5996) - the longest
    decimal prefix of the token, correctly rounded to double then to float; "nan" reads as 0
    and "inf"/"infinity" as +-DBL_MAX (tryParseNanInf, :5906-5948); a non-zero value that
    rounds to 0 or to infinity in double is a parse error (fast_float's out-of-range);
  * indices: sr_parseIntNoSkip (:6752-6777) and fixIndex (:5834-5867);
  * faces: sr_parseTriple (:6789-6836) and the face loop of LoadObj (tiny_obj_loader.h
    :8466-8504); triangulation in exportGroupsToShape (:7080-7600): triangles as they are,
    quads split along the shorter diagonal, larger polygons by tinyobj's built-in ear
    clipping, all in float arithmetic;
  * materials: LoadMtlInternal (:7709-8140), InitMaterial (:6980-7030);
  * the .meta file: LoadMetadata's stream extraction (obj_scene_loader.cc:128-186), whose
    float reads round the decimal text straight to float.

Input the product's loader rejects on purpose (indices beyond the elements defined so far,
which make the reference read out of bounds or drop the face, and malformed MTL files, which
the reference half-loads) raises ObjError here too.
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field

import numpy as np

F32 = np.float32
_EPS = F32(np.finfo(np.float32).eps)
_DECIMAL = re.compile(r"[+-]?(?:[0-9]+(?:\.[0-9]*)?|\.[0-9]+)(?:[eE][+-]?[0-9]+)?")
_DBL_MAX = np.finfo(np.float64).max


class ObjError(Exception):
    pass


# ---------------------------------------------------------------- numbers and tokens
def _token(s: str, i: int) -> tuple[str, int]:
    """Skips blanks, then returns the characters up to the next blank and the new index
    (StreamReader::skip_space then the token loop of sr_parseReal, tiny_obj_loader.h:6627)."""
    n = len(s)
    while i < n and s[i] in " \t":
        i += 1
    j = i
    while j < n and s[j] not in " \t":
        j += 1
    return s[i:j], j


def parse_double(tok: str):
    """tryParseDouble: returns a Python float, or None when the token is not a number."""
    body = tok[1:] if tok[:1] in "+-" else tok
    if body[:1].lower() in ("n", "i"):
        low = body.lower()
        if low.startswith("nan"):
            return 0.0
        if low.startswith("inf"):
            return -_DBL_MAX if tok[0] == "-" else _DBL_MAX
    m = _DECIMAL.match(tok)
    if not m:
        return None
    text = m.group(0)
    val = float(text)
    if math.isinf(val):
        return None
    mant = text.split("e")[0].split("E")[0]
    if val == 0.0 and any(c in "123456789" for c in mant):
        return None
    return val


def to_f32(val: float) -> np.float32:
    with np.errstate(over="ignore"):
        return F32(val)


def parse_real(s: str, i: int, default: float = 0.0):
    """sr_parseReal with a default: an empty token gives the default; a bad token is an
    error (returns ok=False)."""
    tok, j = _token(s, i)
    if tok == "":
        return True, to_f32(default), j
    v = parse_double(tok)
    if v is None:
        return False, to_f32(default), j
    return True, to_f32(v), j


def parse_int_prefix(s: str, i: int) -> tuple[int, int]:
    """sr_parseIntNoSkip: [+-]digits at i; 0 when absent, longer than 63 characters or
    outside int32."""
    n = len(s)
    j = i
    if j < n and s[j] in "+-":
        j += 1
    while j < n and s[j].isdigit() and s[j] in "0123456789":
        j += 1
    text = s[i:j]
    if len(text) > 63:
        return 0, j
    try:
        v = int(text)
    except ValueError:
        return 0, j
    if v > 2**31 - 1 or v < -(2**31):
        return 0, j
    return v, j


def fix_index(idx: int, count: int, allow_zero: bool):
    """fixIndex: 1-based positive, negative relative to the elements read so far. Returns the
    0-based index (or -1), or None for an error."""
    if idx > 0:
        return idx - 1
    if idx == 0:
        return -1 if allow_zero else None
    r = count + idx
    return r if r >= 0 else None


# ---------------------------------------------------------------- materials
@dataclass
class Material:
    name: str = ""
    ambient: list = field(default_factory=lambda: [F32(0)] * 3)
    diffuse: list = field(default_factory=lambda: [F32(0)] * 3)
    specular: list = field(default_factory=lambda: [F32(0)] * 3)
    transmittance: list = field(default_factory=lambda: [F32(0)] * 3)
    emission: list = field(default_factory=lambda: [F32(0)] * 3)
    shininess: np.float32 = F32(1.0)
    ior: np.float32 = F32(1.0)
    dissolve: np.float32 = F32(1.0)

    def row(self) -> np.ndarray:
        """ambient3 diffuse3 specular3 transmittance3 emission3 shininess ior r_ior dissolve,
        with r_ior = 1 / ior as LoadObj sets it (obj_scene_loader.cc:117)."""
        with np.errstate(divide="ignore"):
            r_ior = F32(1.0) / self.ior
        return np.array(list(self.ambient) + list(self.diffuse) + list(self.specular) + list(self.transmittance) +
                        list(self.emission) + [self.shininess, self.ior, r_ior, self.dissolve], np.float32)


def _lines(text: str) -> list[str]:
    if text.startswith("\xef\xbb\xbf"):  # UTF-8 byte order mark, read as latin-1
        text = text[3:]
    out = []
    for ln in re.split(r"\r\n|\r|\n", text):
        z = ln.find("\0")
        out.append(ln if z < 0 else ln[:z])
    return out


def _read(path: str) -> str:
    try:
        with open(path, "r", encoding="latin-1", newline="") as f:
            return f.read()
    except OSError as e:
        raise ObjError(f"cannot read {path}: {e}") from e


def _keyword(ln: str, kw: str) -> bool:
    return ln.startswith(kw) and len(ln) > len(kw) and ln[len(kw)] in " \t"


def load_mtl(path: str, materials: list, material_map: dict) -> None:
    """LoadMtlInternal for the keys a Material holds; appends to materials and material_map
    (the first definition of a name keeps the map entry, as std::map::insert does)."""
    text = _read(path)
    mat = Material()
    has_d = has_kd = False

    def flush():
        if mat.name:
            material_map.setdefault(mat.name, len(materials))
            materials.append(mat)

    def triple(rest):
        out = []
        i = 0
        for _ in range(3):
            ok, v, i = parse_real(rest, i)
            if not ok:
                raise ObjError(f"{path}: expected a number")
            out.append(v)
        return out

    def one(rest):
        ok, v, _ = parse_real(rest, 0)
        if not ok:
            raise ObjError(f"{path}: expected a number")
        return v

    for raw in _lines(text):
        ln = raw.lstrip(" \t")
        if not ln or ln[0] == "#":
            continue
        if _keyword(ln, "newmtl"):
            flush()
            mat = Material()
            has_d = has_kd = False
            mat.name, _ = _token(ln, 7)
        elif _keyword(ln, "Ka"):
            mat.ambient = triple(ln[2:])
        elif _keyword(ln, "Kd"):
            mat.diffuse = triple(ln[2:])
            has_kd = True
        elif _keyword(ln, "Ks"):
            mat.specular = triple(ln[2:])
        elif _keyword(ln, "Kt") or _keyword(ln, "Tf"):
            mat.transmittance = triple(ln[2:])
        elif _keyword(ln, "Ke"):
            mat.emission = triple(ln[2:])
        elif _keyword(ln, "Ni"):
            mat.ior = one(ln[2:])
        elif _keyword(ln, "Ns"):
            mat.shininess = one(ln[2:])
        elif _keyword(ln, "d"):
            mat.dissolve = one(ln[1:])
            has_d = True
        elif _keyword(ln, "Tr"):
            if not has_d:
                mat.dissolve = F32(1.0) - one(ln[2:])
        elif _keyword(ln, "map_Kd"):
            if not has_kd:
                mat.diffuse = [F32(0.6)] * 3
    flush()


def _split_mtllib(rest: str) -> list[str]:
    """SplitString(rest, ' ', '\\\\') without empty tokens: a backslash before a blank or a
    backslash makes that character literal."""
    names, tok, i = [], "", 0
    while i < len(rest):
        c = rest[i]
        if c == "\\" and i + 1 < len(rest) and rest[i + 1] in " \\":
            tok += rest[i + 1]
            i += 2
            continue
        if c == " ":
            if tok:
                names.append(tok)
            tok = ""
        else:
            tok += c
        i += 1
    if tok:
        names.append(tok)
    return names


# ---------------------------------------------------------------- triangulation
def _pnpoly(vx, vy, tx, ty) -> bool:
    c = False
    j = 2
    for i in range(3):
        if (vy[i] > ty) != (vy[j] > ty):
            if tx < (vx[j] - vx[i]) * (ty - vy[i]) / (vy[j] - vy[i]) + vx[i]:
                c = not c
        j = i
    return c


def triangulate(face: list, pos: np.ndarray) -> list:
    """exportGroupsToShape's polygon split. face: list of (v, vt, vn) 0-based triples; pos:
    (count, 3) float32 positions. Returns a list of 3-tuples of face entries."""
    n = len(face)
    if n < 3:
        return []
    if n == 3:
        return [tuple(face)]
    P = lambda e: pos[e[0]]  # noqa: E731
    if n == 4:
        a, b, c, d = (P(e) for e in face)
        e02 = c - a
        e13 = d - b
        s02 = e02[0] * e02[0] + e02[1] * e02[1] + e02[2] * e02[2]
        s13 = e13[0] * e13[0] + e13[1] * e13[1] + e13[2] * e13[2]
        if s02 < s13:
            return [(face[0], face[1], face[2]), (face[0], face[2], face[3])]
        return [(face[0], face[1], face[3]), (face[1], face[2], face[3])]
    # the projection plane: the first corner with a non-degenerate cross product
    axes = [1, 2]
    for k in range(n):
        p0, p1, p2 = P(face[k % n]), P(face[(k + 1) % n]), P(face[(k + 2) % n])
        e0 = p1 - p0
        e1 = p2 - p1
        cx = abs(e0[1] * e1[2] - e0[2] * e1[1])
        cy = abs(e0[2] * e1[0] - e0[0] * e1[2])
        cz = abs(e0[0] * e1[1] - e0[1] * e1[0])
        if cx > _EPS or cy > _EPS or cz > _EPS:
            if not (cx > cy and cx > cz):
                axes[0] = 0
                if cz > cx and cz > cy:
                    axes[1] = 1
            break
    rem = list(face)
    out = []
    guess = 0
    iters = n
    prev = len(rem)
    while len(rem) > 3 and iters > 0:
        m = len(rem)
        if guess >= m:
            guess -= m
        if prev != m:
            prev = m
            iters = m
        else:
            iters -= 1
        ind = [rem[(guess + k) % m] for k in range(3)]
        vx = [P(e)[axes[0]] for e in ind]
        vy = [P(e)[axes[1]] for e in ind]
        e0x = vx[1] - vx[0]
        e0y = vy[1] - vy[0]
        e1x = vx[2] - vx[1]
        e1y = vy[2] - vy[1]
        cross = e0x * e1y - e0y * e1x
        area = (vx[0] * vy[1] - vy[0] * vx[1]) * F32(0.5)
        if cross * area < F32(0.0):
            guess += 1
            continue
        overlap = False
        for other in range(3, m):
            t = P(rem[(guess + other) % m])
            if _pnpoly(vx, vy, t[axes[0]], t[axes[1]]):
                overlap = True
                break
        if overlap:
            guess += 1
            continue
        out.append(tuple(ind))
        del rem[(guess + 1) % m]
    if len(rem) == 3:
        out.append(tuple(rem))
    return out


# ---------------------------------------------------------------- OBJ and scene
@dataclass
class LoadedScene:
    tris: np.ndarray            # (n, 9) float32 positions
    normals: np.ndarray         # (n, 9) float32 vertex normals (zero rows without)
    has_normals: np.ndarray     # (n,) bool
    mat_id: np.ndarray          # (n,) int32, -1 = none
    materials: np.ndarray       # (m, 19) float32, Material.row layout
    material_names: list
    light_tri: np.ndarray       # (k,) int32 area light triangles, in triangle order
    light_area: np.ndarray      # (k,) float32 surface areas
    camera: np.ndarray          # (11,) float32 pos3 dir3 up3 fov far
    point_lights: np.ndarray    # (p, 7) float32 pos3 color3 power


def load_obj(path: str):
    """tinyobj::ObjReader::ParseFromFile followed by LoadObj's conversion."""
    text = _read(path)
    search = os.path.dirname(path)
    v: list = []
    vn: list = []
    vt_count = 0
    faces: list = []  # (material, [(v, vt, vn)])
    materials: list = []
    material_map: dict = {}
    loaded_mtl: set = set()
    material = -1

    def triples(rest: str, what: str):
        out = []
        i, n = 0, len(rest)
        while True:
            while i < n and rest[i] in " \t":
                i += 1
            if i >= n or rest[i] == "#":
                break
            iv, i = parse_int_prefix(rest, i)
            vi = fix_index(iv, len(v), False)
            if vi is None:
                raise ObjError(f"{path}: bad vertex index in `{what}'")
            ti = ni = -1
            while i < n and rest[i] not in "/ \t":
                i += 1
            if i < n and rest[i] == "/":
                i += 1
                if i < n and rest[i] == "/":
                    i += 1
                    x, i = parse_int_prefix(rest, i)
                    ni = fix_index(x, len(vn), True)
                    if ni is None:
                        raise ObjError(f"{path}: bad normal index")
                    while i < n and rest[i] not in "/ \t":
                        i += 1
                else:
                    x, i = parse_int_prefix(rest, i)
                    ti = fix_index(x, vt_count, True)
                    if ti is None:
                        raise ObjError(f"{path}: bad texcoord index")
                    while i < n and rest[i] not in "/ \t":
                        i += 1
                    if i < n and rest[i] == "/":
                        i += 1
                        x, i = parse_int_prefix(rest, i)
                        ni = fix_index(x, len(vn), True)
                        if ni is None:
                            raise ObjError(f"{path}: bad normal index")
                        while i < n and rest[i] not in "/ \t":
                            i += 1
            if vi >= len(v) or ni >= len(vn):
                raise ObjError(f"{path}: index beyond the elements defined so far")
            out.append((vi, ti, ni))
        return out

    for raw in _lines(text):
        ln = raw.lstrip(" \t")
        if not ln or ln[0] == "#":
            continue
        if _keyword(ln, "v"):
            xyz = []
            i = 2
            for _ in range(3):
                ok, val, i = parse_real(ln, i)
                if not ok:
                    raise ObjError(f"{path}: bad vertex")
                xyz.append(val)
            v.append(xyz)
        elif _keyword(ln, "vn"):
            xyz = []
            i = 3
            for _ in range(3):
                ok, val, i = parse_real(ln, i)
                if not ok:
                    raise ObjError(f"{path}: bad normal")
                xyz.append(val)
            vn.append(xyz)
        elif _keyword(ln, "vt"):
            i = 3
            for _ in range(2):
                ok, _val, i = parse_real(ln, i)
                if not ok:
                    raise ObjError(f"{path}: bad texcoord")
            vt_count += 1
        elif _keyword(ln, "f"):
            faces.append((material, triples(ln[2:], "f")))
        elif _keyword(ln, "usemtl"):
            name, _ = _token(ln, 7)
            material = material_map.get(name, -1)
        elif _keyword(ln, "mtllib"):
            for fn in _split_mtllib(ln[7:].rstrip(" \t")):
                if fn in loaded_mtl:  # already read: try the next name (as tinyobj does)
                    continue
                p = os.path.join(search, fn) if search else fn
                if os.path.isfile(p):
                    load_mtl(p, materials, material_map)
                    loaded_mtl.add(fn)
                    break
    pos = np.array(v, np.float32).reshape(-1, 3)
    nrm = np.array(vn, np.float32).reshape(-1, 3)
    out_t, out_n, out_h, out_m = [], [], [], []
    for mat, face in faces:
        for tri in triangulate(face, pos):
            out_t.append(np.concatenate([pos[e[0]] for e in tri]))
            has = all(e[2] >= 0 for e in tri)
            out_n.append(np.concatenate([nrm[e[2]] for e in tri]) if has else np.zeros(9, np.float32))
            out_h.append(has)
            out_m.append(mat)
    return (np.array(out_t, np.float32).reshape(-1, 9), np.array(out_n, np.float32).reshape(-1, 9),
            np.array(out_h, bool), np.array(out_m, np.int32), materials)


def load_meta(path: str):
    """LoadMetadata: whitespace-separated tokens; numbers read as float by stream
    extraction (the decimal text rounded straight to float)."""
    toks = _read(path).split()
    cam = np.zeros(11, np.float32)
    lights = []
    i = 0

    def num():
        nonlocal i
        if i >= len(toks):
            raise ObjError(f"{path}: missing number")
        m = _DECIMAL.match(toks[i])
        if not m or m.end() != len(toks[i]):
            raise ObjError(f"{path}: bad number {toks[i]!r}")
        i += 1
        return _decimal_to_f32(m.group(0))

    slots = {"cam_pos": (0, 3), "cam_dir": (3, 3), "cam_up": (6, 3), "cam_fov": (9, 1), "cam_far": (10, 1)}
    while i < len(toks):
        t = toks[i]
        i += 1
        if t in slots:
            o, k = slots[t]
            for j in range(k):
                cam[o + j] = num()
        elif t == "point_light":
            lights.append([num() for _ in range(7)])
        else:
            raise ObjError(f"{path}: unrecognized metadata token {t!r}")
    return cam, np.array(lights, np.float32).reshape(-1, 7)


def _decimal_to_f32(text: str) -> np.float32:
    """The decimal string correctly rounded to float directly (strtof, as stream extraction
    uses), not through a double: exact rational arithmetic, ties to even. Overflow to
    infinity is a read failure."""
    from fractions import Fraction

    neg = text.startswith("-")
    q = Fraction(text.lstrip("+-"))
    if q == 0:
        return F32(-0.0) if neg else F32(0.0)
    c = to_f32(float(q))  # within one ulp of the answer (double rounding)
    if not np.isfinite(c):
        c = F32(np.finfo(np.float32).max)
    best = None
    for cand in (np.nextafter(c, F32(0)), c, np.nextafter(c, F32(np.inf))):
        if not np.isfinite(cand):
            # rounds to infinity when q is at least the largest float plus half an ulp
            if q >= Fraction(float(np.finfo(np.float32).max)) + Fraction(2) ** 103:
                raise ObjError(f"number out of range: {text}")
            continue
        key = (abs(Fraction(float(cand)) - q), int(np.array(cand, np.float32).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, cand)
    return F32(-best[1]) if neg else F32(best[1])


def _assemble(obj_path: str, cam: np.ndarray, pls: np.ndarray) -> LoadedScene:
    """LoadObj's conversion: triangles in face order, then an area light for every triangle
    whose material emits (obj_scene_loader.cc:36-105)."""
    tris, nrm, has, mat, materials = load_obj(obj_path)
    mrows = np.array([m.row() for m in materials], np.float32).reshape(-1, 19)
    light_tri, light_area = [], []
    for i in range(len(tris)):
        if mat[i] >= 0:
            e = mrows[mat[i], 12:15]
            if (e != 0).any():
                t = tris[i].reshape(3, 3)
                u = t[1] - t[0]
                w = t[2] - t[0]
                c = np.array([u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]],
                             np.float32)
                length = np.sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2])
                light_tri.append(i)
                light_area.append(F32(0.5) * length)
    return LoadedScene(tris, nrm, has, mat, mrows, [m.name for m in materials], np.array(light_tri, np.int32),
                       np.array(light_area, np.float32), cam, pls)


def load_obj_scene(obj_path: str) -> LoadedScene:
    """LoadObj alone (no .meta): zero camera, no point lights."""
    return _assemble(obj_path, np.zeros(11, np.float32), np.zeros((0, 7), np.float32))


def load_scene(dirpath: str, name: str) -> LoadedScene:
    """ObjSceneLoader::Load: the .meta file first, then the .obj (obj_scene_loader.cc:188-204)."""
    base = os.path.join(dirpath, name)
    cam, pls = load_meta(base + ".meta")
    return _assemble(base + ".obj", cam, pls)
