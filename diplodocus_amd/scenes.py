"""Seeded procedural scenes and ray sets for the engine's workloads.

Every generator computes in float64 and rounds to float32 once, so the buffers are the
same on every x86-64 host with the same numpy; tests/golden/ records their hashes.

Workloads (triangle counts are exact):
  * ``cornell_room()``        36 tris: unit-cube room (5 walls), two boxes, light quad.
  * ``icosphere_room()``      36 + 69,632 tris: the room plus a noise-displaced geodesic
                               icosphere (frequency 59 = 69,620 tris, plus centroid splits
                               of 6 triangles = +12), 3-octave value noise, amplitude 0.15.
  * ``mixed_2m()``            2,000,002 tris: 1.8M small tris in 64 anisotropic Gaussian
                               clusters, 200k thin 50:1 tris, 2 ground tris.
  * ``batched_meshes()``      N meshes, log-uniform tri counts (icospheres or soups).
Scene data follow the reference data model (src/scene/scene.h:13-38; triangle.h:12-20;
material.h:7-18; light.h:15-20; camera.h:7-13): positions, per-triangle material ids,
per-material albedo (``Kd``) and emission, camera (pos, dir, up, fov, far).
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "Camera",
    "SceneData",
    "cornell_room",
    "icosphere_room",
    "mixed_2m",
    "random_rays_towards",
    "camera_rays",
    "batched_meshes",
    "buffer_hash",
]


@dataclass
class Camera:
    """src/scene/camera.h:7-13 (fov in radians, far = ray t_max)."""

    pos: tuple = (0.5, 0.5, -1.35)
    dir: tuple = (0.0, 0.0, 1.0)
    up: tuple = (0.0, 1.0, 0.0)
    fov: float = 0.75
    far: float = 100.0

    def as_array(self) -> np.ndarray:
        return np.array([*self.pos, *self.dir, *self.up, self.fov, self.far], dtype=np.float32)


@dataclass
class SceneData:
    tris: np.ndarray  # (n, 9) float32: v0 xyz, v1 xyz, v2 xyz
    mat_id: np.ndarray  # (n,) int32
    albedo: np.ndarray  # (M, 3) float32
    emission: np.ndarray  # (M, 3) float32
    camera: Camera = field(default_factory=Camera)

    @property
    def n(self) -> int:
        return int(self.tris.shape[0])


def buffer_hash(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


# --------------------------------------------------------------------------------------
# Cornell-style room (36 tris)
# --------------------------------------------------------------------------------------
def _quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3); counter-clockwise seen from the front."""
    return [(p0, p1, p2), (p0, p2, p3)]


def _box(lo, hi):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    q = []
    q += _quad((x0, y1, z0), (x0, y1, z1), (x1, y1, z1), (x1, y1, z0))  # top (+y)
    q += _quad((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1))  # bottom (-y)
    q += _quad((x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (x1, y0, z0))  # front (-z)
    q += _quad((x1, y0, z1), (x1, y1, z1), (x0, y1, z1), (x0, y0, z1))  # back (+z)
    q += _quad((x0, y0, z1), (x0, y1, z1), (x0, y1, z0), (x0, y0, z0))  # left (-x)
    q += _quad((x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1))  # right (+x)
    return q


def _room_tris():
    """(tris, group) — groups: 0 floor, 1 ceiling, 2 back, 3 left, 4 right, 5 short box,
    6 tall box, 7 light."""
    t, g = [], []

    def add(lst, grp):
        t.extend(lst)
        g.extend([grp] * len(lst))

    add(_quad((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0)), 0)  # floor, normal +y
    add(_quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)), 1)  # ceiling, normal -y
    add(_quad((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1)), 2)  # back wall, normal -z
    add(_quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)), 3)  # left wall, normal +x
    add(_quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)), 4)  # right wall, normal -x
    add(_box((0.55, 0.0, 0.15), (0.85, 0.3, 0.45)), 5)
    add(_box((0.15, 0.0, 0.5), (0.45, 0.6, 0.8)), 6)
    add(_quad((0.35, 0.99, 0.35), (0.65, 0.99, 0.35), (0.65, 0.99, 0.65), (0.35, 0.99, 0.65)), 7)  # light, -y
    tris = np.array(t, dtype=np.float64).reshape(-1, 9)
    return tris, np.array(g, dtype=np.int32)


def cornell_room() -> SceneData:
    """36 tris in the unit cube: 5 walls (10), two boxes (24), light quad (2)."""
    tris, grp = _room_tris()
    albedo = np.array(
        [[0.73, 0.73, 0.73], [0.73, 0.73, 0.73], [0.73, 0.73, 0.73], [0.63, 0.065, 0.05], [0.14, 0.45, 0.091],
         [0.73, 0.73, 0.73], [0.73, 0.73, 0.73], [0.78, 0.78, 0.78]], dtype=np.float32)
    emission = np.zeros((8, 3), dtype=np.float32)
    emission[7] = (17.0, 12.0, 4.0)
    return SceneData(tris.astype(np.float32), grp, albedo, emission, Camera())


def whitted_material(diffuse, specular=(0.0, 0.0, 0.0), transmittance=(0.0, 0.0, 0.0), emission=(0.0, 0.0, 0.0),
                     shininess=1.0, ior=1.0) -> np.ndarray:
    """One Whitted material row (16 float32): diffuse3 specular3 transmittance3 emission3
    shininess ior r_ior pad, r_ior = 1 / ior (src/io/scene/obj_scene_loader.cc:118)."""
    m = np.zeros(16, np.float32)
    m[0:3], m[3:6], m[6:9], m[9:12] = diffuse, specular, transmittance, emission
    m[12], m[13] = shininess, ior
    m[14] = np.float32(1.0) / np.float32(ior)
    return m


def cornell_whitted():
    """The 36-triangle room with Whitted materials: white floor / ceiling / back wall, red
    left and green right walls, a glossy short box (Phong + mirror), a glass tall box
    (refraction, ior 1.5) and the emissive light quad. Returns (scene, materials (8, 16))."""
    s = cornell_room()
    mats = np.stack([
        whitted_material((0.73, 0.73, 0.73)), whitted_material((0.73, 0.73, 0.73)),
        whitted_material((0.73, 0.73, 0.73)), whitted_material((0.63, 0.065, 0.05)),
        whitted_material((0.14, 0.45, 0.091)),
        whitted_material((0.3, 0.3, 0.35), specular=(0.6, 0.6, 0.6), shininess=64.0),
        whitted_material((0.05, 0.05, 0.05), specular=(0.1, 0.1, 0.1), transmittance=(0.85, 0.85, 0.85),
                         shininess=128.0, ior=1.5),
        whitted_material((0.0, 0.0, 0.0), emission=(17.0, 12.0, 4.0)),
    ])
    return s, mats


def sphere_vertex_normals(tris: np.ndarray, first: int, center) -> np.ndarray:
    """Per-vertex normals (n, 9): unit (vertex - center) for triangles first.., zero (none)
    before."""
    vn = np.zeros_like(tris, dtype=np.float32)
    v = tris[first:].reshape(-1, 3, 3).astype(np.float64) - np.asarray(center, np.float64)
    v /= np.linalg.norm(v, axis=2, keepdims=True)
    vn[first:] = v.reshape(-1, 9).astype(np.float32)
    return vn


# --------------------------------------------------------------------------------------
# Value noise and the displaced geodesic icosphere
# --------------------------------------------------------------------------------------
def _hash_u32(x: np.ndarray) -> np.ndarray:
    """The reference's avalanche hash (src/util/util.h:24-32), vectorised on uint32."""
    x = x.astype(np.uint32)
    x ^= x >> np.uint32(16)
    x *= np.uint32(0x7FEB352D)
    x ^= x >> np.uint32(15)
    x *= np.uint32(0x846CA68B)
    x ^= x >> np.uint32(16)
    return x


def _lattice(ix, iy, iz, seed):
    k = (ix.astype(np.int64) * 73856093) ^ (iy.astype(np.int64) * 19349663) ^ (iz.astype(np.int64) * 83492791)
    k = (k ^ seed) & 0xFFFFFFFF
    return _hash_u32(k.astype(np.uint32)).astype(np.float64) / 4294967295.0 * 2.0 - 1.0


def value_noise3(p: np.ndarray, seed: int, octaves: int = 3) -> np.ndarray:
    """Sum of `octaves` trilinear value-noise octaves (frequency x2, amplitude x0.5),
    smoothstep-interpolated, normalised to [-1, 1]."""
    total = np.zeros(p.shape[0])
    amp, freq, norm = 1.0, 1.0, 0.0
    for o in range(octaves):
        q = p * freq
        f = np.floor(q)
        i = f.astype(np.int64)
        t = q - f
        s = t * t * (3.0 - 2.0 * t)
        acc = np.zeros(p.shape[0])
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    w = ((s[:, 0] if dx else 1 - s[:, 0]) * (s[:, 1] if dy else 1 - s[:, 1]) *
                         (s[:, 2] if dz else 1 - s[:, 2]))
                    acc += w * _lattice(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed + 7919 * o)
        total += amp * acc
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def geodesic_sphere(freq: int):
    """Class-I geodesic subdivision of the icosahedron: 20 * freq^2 outward-facing tris.

    Vertices are keyed by their integer barycentric weights over the sorted icosahedron
    corners, so vertices shared by neighbouring faces are computed once (watertight)."""
    V, F = _icosahedron()
    key_index = {}
    pos = []
    tris = []

    def vid(face, i, j):
        w = {face[0]: freq - i - j, face[1]: i, face[2]: j}
        key = tuple(sorted((c, wt) for c, wt in w.items() if wt))
        k = key_index.get(key)
        if k is None:
            p = sum(V[c] * wt for c, wt in key) / freq
            k = len(pos)
            key_index[key] = k
            pos.append(p / np.linalg.norm(p))
        return k

    for face in F:
        face = [int(x) for x in face]
        for i in range(freq):
            for j in range(freq - i):
                tris.append((vid(face, i, j), vid(face, i + 1, j), vid(face, i, j + 1)))
                if i + j < freq - 1:
                    tris.append((vid(face, i + 1, j), vid(face, i + 1, j + 1), vid(face, i, j + 1)))
    return np.array(pos), np.array(tris, dtype=np.int64)


def displaced_icosphere(freq: int, splits: int, center, radius: float, amplitude: float, seed: int,
                        noise_freq: float = 4.0):
    """Noise-displaced geodesic sphere; `splits` triangles (every 997th) are split at their
    centroid into three, adding 2*splits triangles, keeping the mesh closed."""
    P, T = geodesic_sphere(freq)
    if splits:
        pick = (np.arange(splits) * 997) % T.shape[0]
        extra_v = []
        new_t = [T[np.setdiff1d(np.arange(T.shape[0]), pick)]]
        for k, ti in enumerate(pick):
            a, b, c = T[ti]
            m = P[a] + P[b] + P[c]
            extra_v.append(m / np.linalg.norm(m))
            mi = P.shape[0] + k
            new_t.append(np.array([[a, b, mi], [b, c, mi], [c, a, mi]]))
        P = np.vstack([P, np.array(extra_v)])
        T = np.vstack(new_t)
    r = radius * (1.0 + amplitude * value_noise3(P * noise_freq + 17.0, seed))
    X = np.asarray(center, dtype=np.float64) + P * r[:, None]
    return X[T].reshape(-1, 9)


def icosphere_room(seed: int = 7, n_materials: int = 16) -> SceneData:
    """Room (36 tris) + 69,632-tri displaced icosphere. With ``n_materials=16`` the
    materials are: 0-4 walls, 5-6 boxes, 7-14 sphere octants, 15 the emissive light;
    albedo U[0.1, 0.9] per channel (seeded)."""
    room, grp = _room_tris()
    sph = displaced_icosphere(59, 6, center=(0.68, 0.6, 0.45), radius=0.16, amplitude=0.15, seed=seed)
    tris = np.vstack([room, sph]).astype(np.float32)
    rng = np.random.default_rng(seed)
    albedo = rng.uniform(0.1, 0.9, size=(n_materials, 3)).astype(np.float32)
    emission = np.zeros((n_materials, 3), dtype=np.float32)
    emission[n_materials - 1] = (12.0, 12.0, 12.0)
    mat = np.empty(tris.shape[0], dtype=np.int32)
    room_map = np.array([0, 1, 2, 3, 4, 5, 6, n_materials - 1], dtype=np.int32)
    mat[: room.shape[0]] = room_map[grp]
    c = sph.reshape(-1, 3, 3).mean(axis=1) - np.array([0.68, 0.6, 0.45])
    octant = (c[:, 0] > 0).astype(np.int32) + 2 * (c[:, 1] > 0) + 4 * (c[:, 2] > 0)
    mat[room.shape[0]:] = 7 + octant
    return SceneData(tris, mat, albedo, emission, Camera())


# --------------------------------------------------------------------------------------
# 2M-triangle mixed-distribution scene
# --------------------------------------------------------------------------------------
def _random_rotations(rng, k):
    q = rng.normal(size=(k, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)


def _unit_vectors(rng, k):
    v = rng.normal(size=(k, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def mixed_2m(seed: int = 3, n_small: int = 1_800_000, n_thin: int = 200_000, n_clusters: int = 64) -> SceneData:
    """64 anisotropic Gaussian clusters (sigma ratio up to 20:1, random rotations) of small
    triangles, long thin triangles (aspect 50:1, random axes) and 2 ground triangles."""
    rng = np.random.default_rng(seed)
    lo, hi = np.array([-50.0, 0.0, -50.0]), np.array([50.0, 40.0, 50.0])
    centers = rng.uniform(lo, hi, size=(n_clusters, 3))
    major = rng.uniform(2.0, 8.0, size=n_clusters)
    ratios = np.stack([np.ones(n_clusters), rng.uniform(1 / 20, 1.0, n_clusters), np.full(n_clusters, 1 / 20)], 1)
    ratios[:, 1:] = np.sort(ratios[:, 1:], axis=1)[:, ::-1]
    sig = major[:, None] * ratios
    rot = _random_rotations(rng, n_clusters)
    which = rng.integers(0, n_clusters, size=n_small)
    local = rng.normal(size=(n_small, 3)) * sig[which]
    cen = centers[which] + np.einsum("nij,nj->ni", rot[which], local)
    size = rng.uniform(0.03, 0.15, size=n_small)
    small = np.concatenate([cen + _unit_vectors(rng, n_small) * size[:, None] for _ in range(3)], axis=1)

    length = rng.uniform(2.0, 12.0, size=n_thin)
    axis = _unit_vectors(rng, n_thin)
    side = np.cross(axis, _unit_vectors(rng, n_thin))
    side /= np.linalg.norm(side, axis=1, keepdims=True)
    mid = rng.uniform(lo, hi, size=(n_thin, 3))
    half = (length / 2)[:, None]
    width = (length / 50)[:, None]
    thin = np.concatenate([mid - axis * half, mid + axis * half, mid + side * width], axis=1)

    g = 120.0
    ground = np.array([[-g, -1.0, -g, -g, -1.0, g, g, -1.0, g], [-g, -1.0, -g, g, -1.0, g, g, -1.0, -g]])
    tris = np.vstack([small, thin, ground]).astype(np.float32)
    n = tris.shape[0]
    return SceneData(tris, np.zeros(n, np.int32), np.full((1, 3), 0.5, np.float32), np.zeros((1, 3), np.float32),
                     Camera(pos=(0.0, 20.0, -150.0), dir=(0.0, 0.0, 1.0), up=(0.0, 1.0, 0.0), fov=0.9, far=1e4))


def random_rays_towards(tris: np.ndarray, n: int, seed: int = 11, jitter: float = 0.2) -> np.ndarray:
    """Rays (n, 8) float32 = (o xyz, tmin, d xyz, tmax): origins uniform on a sphere of
    1.2 x the scene's half-diagonal around the scene centre, aimed at the centre displaced
    uniformly by +-jitter x the scene extent per axis; tmin 1e-6, tmax FLT_MAX."""
    rng = np.random.default_rng(seed)
    v = tris.reshape(-1, 3).astype(np.float64)
    lo, hi = v.min(0), v.max(0)
    c = (lo + hi) / 2
    rad = 1.2 * np.linalg.norm(hi - lo) / 2
    o = c + _unit_vectors(rng, n) * rad
    tgt = c + rng.uniform(-1, 1, size=(n, 3)) * (hi - lo) * jitter
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.empty((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 1e-6
    r[:, 4:7] = d
    r[:, 7] = np.finfo(np.float32).max
    return r


def batched_rays(tris: np.ndarray, offsets: np.ndarray, per_mesh: int, seed: int = 13, jitter: float = 0.2):
    """random_rays_towards for every mesh of a batch at once: per_mesh rays per mesh from a
    sphere around that mesh, aimed at its centre. Returns (rays (M*per_mesh, 8) float32,
    ray_mesh (M*per_mesh,) int32), mesh-major."""
    rng = np.random.default_rng(seed)
    m = len(offsets) - 1
    v = tris.reshape(-1, 9)
    lo = np.minimum.reduceat(np.minimum(np.minimum(v[:, 0:3], v[:, 3:6]), v[:, 6:9]), offsets[:-1]).astype(np.float64)
    hi = np.maximum.reduceat(np.maximum(np.maximum(v[:, 0:3], v[:, 3:6]), v[:, 6:9]), offsets[:-1]).astype(np.float64)
    ids = np.repeat(np.arange(m, dtype=np.int32), per_mesh)
    c = ((lo + hi) / 2)[ids]
    ext = (hi - lo)[ids]
    rad = 1.2 * np.linalg.norm(ext, axis=1, keepdims=True) / 2
    o = c + _unit_vectors(rng, len(ids)) * rad
    d = c + rng.uniform(-1, 1, size=(len(ids), 3)) * ext * jitter - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.empty((len(ids), 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 1e-6
    r[:, 4:7] = d
    r[:, 7] = np.finfo(np.float32).max
    return r, ids


def camera_rays(cam: Camera, w: int, h: int) -> np.ndarray:
    """One primary ray per pixel, row-major, following src/renderer/cpu_raytracer.cc:62-73
    and :99-106 (no pixel-centre offset), evaluated in float32 on the host."""
    f = np.float32
    pos = np.array(cam.pos, f)
    d = np.array(cam.dir, f)
    up = np.array(cam.up, f)

    def normalize(v):
        ln = np.sqrt((v * v).sum(-1, keepdims=True, dtype=f)).astype(f)
        return (v / ln).astype(f)

    right = normalize(np.cross(d, up).astype(f))
    gw = f(2.0 * np.tan(f(cam.fov) / f(2.0)))
    gh = f(gw * (f(h) / f(w)))
    qw = (right * f(gw / f(w - 1))).astype(f)
    qh = (-up * f(gh / f(h - 1))).astype(f)
    p00 = (d - right * f(gw / f(2)) + up * f(gh / f(2))).astype(f)
    ys, xs = np.meshgrid(np.arange(h, dtype=f), np.arange(w, dtype=f), indexing="ij")
    v = p00[None, :] + qw[None, :] * xs.reshape(-1, 1) + qh[None, :] * ys.reshape(-1, 1)
    dirs = normalize(v.astype(f))
    r = np.empty((w * h, 8), f)
    r[:, 0:3] = pos
    r[:, 3] = 1e-6
    r[:, 4:7] = dirs
    r[:, 7] = cam.far
    return r


def batched_meshes(count: int = 4096, lo: int = 64, hi: int = 65536, seed: int = 5):
    """`count` independent meshes with log-uniform triangle counts in [lo, hi]: even
    indices random-triangle soups in a unit cube, odd indices noise-displaced spheres
    (geodesic frequency chosen to approach the count, then trimmed to it). Returns
    (tris (N,9) float32, offsets (count+1,) int64)."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(lo), np.log(hi), size=count)).astype(np.int64)
    parts = []
    sphere_cache = {}
    for m, k in enumerate(counts):
        k = int(k)
        if m % 2 == 0:
            c = rng.uniform(0.2, 0.8, size=(k, 1, 3))
            t = c + rng.uniform(-0.05, 0.05, size=(k, 3, 3))
            parts.append(t.reshape(k, 9))
        else:
            fq = max(2, int(np.ceil(np.sqrt(k / 20.0))))
            if fq not in sphere_cache:
                P, T = geodesic_sphere(fq)
                sphere_cache[fq] = (P, T)
            P, T = sphere_cache[fq]
            r = 0.3 * (1.0 + 0.15 * value_noise3(P * 4.0 + m, seed + m))
            X = 0.5 + P * r[:, None]
            parts.append(X[T].reshape(-1, 9)[:k])
    offsets = np.zeros(count + 1, np.int64)
    offsets[1:] = np.cumsum([p.shape[0] for p in parts])
    return np.vstack(parts).astype(np.float32), offsets
