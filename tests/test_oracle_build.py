"""Pins the oracle's restatement of the engine build (codes, order, PLOC, collapse, final
layout). No reference implementation exists (src/gpu/acceleration/ is empty), so these are
known answers from the definitions, invariants, and brute force on small inputs."""
import numpy as np
import pytest

from diplodocus_amd import scenes


def tri_at(c, s=0.01):
    c = np.asarray(c, np.float32)
    return np.concatenate([c + [-s, -s, 0], c + [s, -s, 0], c + [0, s, 0]]).astype(np.float32)


def soup(n, seed, spread=10.0, size=0.2):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, spread, (n, 1, 3))
    return (c + rng.normal(0, size, (n, 3, 3))).reshape(n, 9).astype(np.float32)


# ---------------------------------------------------------------------------------------
# Morton30
# ---------------------------------------------------------------------------------------
def test_expand_bits_known_answers(oracle):
    f = oracle.lib.orc_expand_bits10
    assert f(0) == 0
    assert f(1) == 1
    assert f(0b11) == 0b1001
    assert f(0x3FF) == 0x09249249
    for v in range(1024):  # bit k of v moves to bit 3k
        e = f(v)
        assert e == sum(((v >> k) & 1) << (3 * k) for k in range(10))


def test_morton30_corners_and_axes(oracle):
    # centroids exactly at the centroid-box corners: (0,0,0) -> 0 and (1,1,1) -> all ones
    tris = np.stack([tri_at([0, 0, 0]), tri_at([1, 1, 1]), tri_at([1, 0, 0]), tri_at([0, 1, 0]), tri_at([0, 0, 1])])
    c = oracle.codes(tris, 0)
    cen = np.array([oracle.centroid(t) for t in tris])
    lo, hi = cen.min(0), cen.max(0)
    assert c[0] == 0
    assert c[1] == 0x3FFFFFFF
    assert c[2] == 0x24924924  # x bits in position 3k+2
    assert c[3] == 0x12492492  # y bits in position 3k+1
    assert c[4] == 0x09249249  # z bits in position 3k
    assert (lo == cen[0]).all() and (hi == cen[1]).all()


def test_morton30_quantisation_steps(oracle):
    # a centroid a fraction f along x quantises to floor(1024 f) (clamped to 1023)
    xs = [0.0, 0.5, 0.25, 1.0 / 1024, 1023.5 / 1024, 1.0]
    tris = np.stack([tri_at([x, 0, 0]) for x in xs] + [tri_at([0, 1, 1])])
    c = oracle.codes(tris, 0)
    cen = np.array([oracle.centroid(t) for t in tris])
    e = oracle.lib.orc_expand_bits10
    for k in range(len(xs)):
        fx = np.float32(cen[k, 0] - cen[:, 0].min()) * np.float32(np.float32(1024) / np.float32(np.ptp(cen[:, 0])))
        q = min(int(fx), 1023)
        assert c[k] == e(q) << 2


# ---------------------------------------------------------------------------------------
# EMC60
# ---------------------------------------------------------------------------------------
def spatial_bits(code):
    """Drops the size bit of every 5-bit group of an EMC60 code -> 48-bit spatial code."""
    out = 0
    for g in range(12):
        grp = (int(code) >> (55 - 5 * g)) & 0x1F
        out = (out << 4) | (grp & 0xF)
    return out


def size_bits(code):
    return sum(((int(code) >> (59 - 5 * g)) & 1) << (11 - g) for g in range(12))


def test_emc_layout_adaptive_allocation(oracle):
    b = np.array([0, 0, 0, 8, 2, 1, 0, 0, 0, 8, 2, 1], np.float32)  # extents 8:2:1
    nb, ax = oracle.emc_layout(b)
    # cells (8,2,1): x,x,x (tie 2=2 goes to x), y, then x,y,z repeating from cells (1,1,1)
    assert list(ax[:7]) == [0, 0, 0, 1, 0, 1, 2]
    assert list(ax[7:]) == ([0, 1, 2] * 14)[:41]
    assert list(nb) == [18, 16, 14]
    cube = np.array([0, 0, 0, 1, 1, 1, 0, 0, 0, 1, 1, 1], np.float32)
    nb, ax = oracle.emc_layout(cube)
    assert list(nb) == [16, 16, 16] and list(ax) == [0, 1, 2] * 16
    flat = np.array([0, 0, 0, 1, 1, 0, 0, 0, 0, 1, 1, 0], np.float32)  # zero z extent
    nb, ax = oracle.emc_layout(flat)
    assert nb[2] == 0 and nb[0] == 24 and nb[1] == 24


def test_emc60_reduces_to_morton_in_a_cube(oracle):
    # With cube centroid bounds the allocation is x,y,z repeating; dropping the size bits
    # gives a 48-bit Morton code whose top 30 bits equal Morton30 (2^16/e = 64 * 2^10/e).
    rng = np.random.default_rng(4)
    cen = rng.uniform(0, 1, (2000, 3))
    cen[0] = 0.0
    cen[1] = 1.0
    tris = np.stack([tri_at(c, 0.003) for c in cen])
    m30 = oracle.codes(tris, 0)
    emc = oracle.codes(tris, 1)
    for a, b in zip(m30, emc):
        assert spatial_bits(b) >> 18 == int(a)


def test_emc60_size_bits_are_monotone_in_size(oracle):
    c = [0.5, 0.5, 0.5]
    tris = np.stack([tri_at(c, s) for s in [1e-5, 1e-4, 1e-3, 1e-2, 0.1, 0.5]] + [tri_at([0, 0, 0]), tri_at([1, 1, 1])])
    emc = oracle.codes(tris, 1)
    sz = [size_bits(e) for e in emc[:6]]
    assert sz == sorted(sz) and sz[0] < sz[-1]
    assert all(0 <= s < 4096 for s in sz)


# ---------------------------------------------------------------------------------------
# stable order, nearest neighbours
# ---------------------------------------------------------------------------------------
def test_sorted_order_is_the_stable_order(oracle):
    tris = soup(3000, 5, size=0.5)
    tris[100:200] = tris[100]  # duplicate keys keep input order
    for mt in (0, 1):
        p = oracle.Ploc(tris, mt, 8, 4)
        assert (p.sorted_idx == np.argsort(p.codes, kind="stable")).all()
        assert (p.sorted_codes == p.codes[p.sorted_idx]).all()


def _pair_hash(oracle, i, j):
    lo, hi = min(i, j), max(i, j)
    return oracle.hash_u32((lo * 0x9E3779B1 ^ hi) & 0xFFFFFFFF)


@pytest.mark.parametrize("radius", [1, 3, 8])
def test_nearest_neighbour_brute_force(oracle, radius):
    # Integer-coordinate boxes: SA(union) is exact in float32 and in int64, so the argmin
    # of (distance, pair priority, index) can be recomputed independently.
    rng = np.random.default_rng(radius)
    n = 300
    lo = rng.integers(0, 6, (n, 3))
    hi = lo + rng.integers(0, 3, (n, 3))
    lohi = np.concatenate([lo, hi], 1).astype(np.float32)
    nn = oracle.ploc_nn(lohi, radius)
    for i in range(n):
        best = None
        for j in range(max(0, i - radius), min(n, i + radius + 1)):
            if j == i:
                continue
            s = np.maximum(hi[i], hi[j]) - np.minimum(lo[i], lo[j])
            d = 2 * (s[0] * s[1] + s[1] * s[2] + s[0] * s[2])
            key = (d, _pair_hash(oracle, i, j), j)
            best = key if best is None or key < best else best
        assert nn[i] == best[2]


@pytest.mark.parametrize("n,radius,seed", [(300, 1, 1), (400, 4, 2), (500, 16, 3), (257, 8, 4)])
def test_merge_list_replays_mutual_nearest_neighbours(oracle, n, radius, seed):
    """Replays the recorded merge list iteration by iteration: the nodes made in iteration k
    (ids 2n - c_k .. 2n - c_k+1 - 1) must be exactly the mutual nearest-neighbour pairs of
    that iteration's cluster list, in list order, each replacing its lower-index member."""
    tris = soup(n, seed, size=0.3)
    p = oracle.Ploc(tris, 0, radius, 4)
    nodes = p.nodes
    clusters = list(range(n))  # node ids, in code order
    ic = list(p.iter_counts)
    for k, c in enumerate(ic):
        assert len(clusters) == c
        lohi = np.concatenate([nodes["lo"][clusters], nodes["hi"][clusters]], 1).astype(np.float32)
        nn = oracle.ploc_nn(lohi, radius)
        pairs = [(i, int(nn[i])) for i in range(c) if nn[nn[i]] == i and i < nn[i]]
        assert pairs, "every iteration merges at least one pair"
        base = 2 * n - c
        nxt, partner = [], {j: i for i, j in pairs}
        made = dict((i, base + r) for r, (i, _) in enumerate(pairs))
        for r, (i, j) in enumerate(pairs):
            nd = nodes[base + r]
            assert nd["left"] == clusters[i] and nd["right"] == clusters[j]
            assert (nd["lo"] == np.minimum(nodes["lo"][clusters[i]], nodes["lo"][clusters[j]])).all()
        for i in range(c):
            if i in made:
                nxt.append(made[i])
            elif i not in partner:
                nxt.append(clusters[i])
        clusters = nxt
        nexpected = ic[k + 1] if k + 1 < len(ic) else 1
        assert len(clusters) == nexpected
    assert clusters == [2 * n - 2]


# ---------------------------------------------------------------------------------------
# PLOC merges, collapse, final layout
# ---------------------------------------------------------------------------------------
def check_ploc_invariants(p, tris, max_leaf):
    n = len(tris)
    nodes, parent = p.nodes, p.parent
    root = 2 * n - 2
    assert parent[root] == -1
    # leaves: node k is the k-th triangle in sorted order, with its exact box
    box = tris.reshape(n, 3, 3)[p.sorted_idx]
    assert (nodes["left"][:n] == p.sorted_idx.astype(np.int32)).all()
    assert (nodes["lo"][:n] == box.min(1)).all() and (nodes["hi"][:n] == box.max(1)).all()
    # internal nodes: children created earlier, each node has exactly one parent
    inner = nodes[n:]
    ids = np.arange(n, 2 * n - 1)
    assert (inner["left"] < ids).all() and (inner["right"] < ids).all()
    kids = np.concatenate([inner["left"], inner["right"]])
    assert len(np.unique(kids)) == 2 * n - 2 and root not in kids
    assert (parent[inner["left"]] == ids).all() and (parent[inner["right"]] == ids).all()
    # box(parent) == union(children), exactly
    L, R = nodes[inner["left"]], nodes[inner["right"]]
    assert (inner["lo"] == np.minimum(L["lo"], R["lo"])).all()
    assert (inner["hi"] == np.maximum(L["hi"], R["hi"])).all()
    # iteration counts strictly decrease and every iteration merges
    ic = p.iter_counts
    if n > 1:
        assert ic[0] == n and (np.diff(ic) < 0).all()
    # collapse: counts, leaf sizes
    assert (p.count[:n] == 1).all() and p.count[root] == n
    # final layout: prim_order is a permutation; leaf ranges partition it
    assert sorted(p.prim_order.tolist()) == list(range(n))
    refs = np.concatenate([p.fnodes["left"], p.fnodes["right"]])
    lr = ~refs[refs < 0]
    lr = lr[(lr & 15) > 0]
    begins, counts = lr >> 4, lr & 15
    assert counts.max() <= max_leaf
    order = np.argsort(begins)
    assert (np.cumsum(counts[order])[:-1] == begins[order][1:]).all() and counts.sum() == n


@pytest.mark.parametrize("n,radius,max_leaf,mt", [(1, 16, 4, 0), (2, 16, 4, 0), (3, 1, 1, 0), (5, 2, 4, 1),
                                                  (200, 4, 4, 0), (2500, 16, 4, 1), (3000, 32, 15, 0),
                                                  (3000, 8, 1, 1)])
def test_ploc_invariants(oracle, n, radius, max_leaf, mt):
    tris = soup(n, n + radius)
    p = oracle.Ploc(tris, mt, radius, max_leaf)
    check_ploc_invariants(p, tris, max_leaf)
    rays = scenes.random_rays_towards(tris, 2000, seed=n)
    assert (p.trace(rays) == oracle.trace_brute(tris, rays)).all()


def test_ploc_identical_triangles_progress(oracle):
    # 4000 copies of one triangle: every distance ties; the pair-priority rule still merges
    # a constant fraction per iteration (a lower-index rule would merge one pair).
    tris = np.repeat(tri_at([1, 2, 3])[None], 4000, 0)
    p = oracle.Ploc(tris, 0, 16, 4)
    assert len(p.iter_counts) < 80
    check_ploc_invariants(p, tris, 4)


def test_collapse_is_cost_reducing(oracle):
    tris = soup(3000, 9)
    p1 = oracle.Ploc(tris, 0, 16, 1)  # max_leaf 1: no collapse
    p4 = oracle.Ploc(tris, 0, 16, 4)
    assert (p1.nodes == p4.nodes).all()
    assert p4.sah() <= p1.sah()
    # the collapse itself is pinned against brute-force enumeration of every leaf cut in
    # tests/test_oracle_pins.py; here: collapsed nodes hold at most max_leaf triangles
    assert (p4.count[p4.collapsed.astype(bool)] <= 4).all()


def test_final_layout_sah_closed_form(oracle):
    # two triangles far apart: root with two single-triangle leaves
    a = tri_at([0, 0, 0], 0.5)
    b = tri_at([10, 0, 0], 0.5)
    p = oracle.Ploc(np.stack([a, b]), 0, 16, 4)
    assert len(p.fnodes) == 1
    box = lambda t: (t.reshape(3, 3).min(0).astype(np.float64), t.reshape(3, 3).max(0).astype(np.float64))

    def sa(lo, hi):
        s = hi - lo
        return 2 * (s[0] * s[1] + s[1] * s[2] + s[0] * s[2])

    la, ha = box(a)
    lb, hb = box(b)
    root = sa(np.minimum(la, lb), np.maximum(ha, hb))
    assert p.sah() == pytest.approx((root + sa(la, ha) + sa(lb, hb)) / root, rel=1e-12)


def test_sobb_direction_set(oracle):
    dirs, abc, inv = oracle.sobb_set(32)
    assert len(dirs) == 13 and np.allclose(np.linalg.norm(dirs.astype(np.float64), axis=1), 1, atol=1e-7)
    assert list(abc[0]) == [0, 1, 2] and inv[0] == 1.0  # the axis-aligned box comes first
    for (a, b, c), iv in zip(abc, inv):
        det = abs(np.linalg.det(dirs[[a, b, c]].astype(np.float64)))
        assert det >= 1e-3 and iv == np.float32(1.0 / det)
    assert len(oracle.sobb_set(6)[1]) == 1 and len(oracle.sobb_set(18)[0]) == 9 and len(oracle.sobb_set(14)[0]) == 7


def test_sobb_kdops_contain_their_triangles_in_fp64(oracle):
    tris = soup(3000, 21, size=0.4)
    p = oracle.Ploc(tris, 1, 16, 4, bv_type=1, kdop_size=32)
    dirs, _, _ = oracle.sobb_set(32)
    nd = len(dirs)
    n = len(tris)
    v = tris.reshape(n, 3, 3).astype(np.float64)[p.sorted_idx]  # leaf k = sorted triangle k
    proj = np.einsum("kd,nvd->nvk", dirs.astype(np.float64), v)  # exact-ish projections
    kd = p.kdop.astype(np.float64)
    assert (proj.min(1) >= kd[:n, :nd]).all() and (proj.max(1) <= kd[:n, nd:]).all()
    # unions are exact: every inner k-DOP is the min/max of its children's
    nodes = p.nodes[n:]
    L, R = p.kdop[nodes["left"]], p.kdop[nodes["right"]]
    assert (p.kdop[n:, :nd] == np.minimum(L[:, :nd], R[:, :nd])).all()
    assert (p.kdop[n:, nd:] == np.maximum(L[:, nd:], R[:, nd:])).all()


def test_sobb_tree_traces_like_brute_force_and_lowers_sah(oracle):
    s = scenes.mixed_2m(n_small=30_000, n_thin=5_000)
    rays = scenes.random_rays_towards(s.tris, 3000, seed=4)
    brute = oracle.trace_brute(s.tris, rays)
    aabb = oracle.Ploc(s.tris, 1, 16, 4)
    for kdop in (32, 18, 14, 6):
        p = oracle.Ploc(s.tris, 1, 16, 4, bv_type=1, kdop_size=kdop)
        assert (p.trace(rays) == brute).all()
        if kdop > 6:
            assert p.sah() < aabb.sah()


@pytest.mark.parametrize("scene", ["cornell", "icosphere"])
def test_ploc_on_workload_scenes(oracle, scene):
    s = scenes.cornell_room() if scene == "cornell" else scenes.icosphere_room()
    for mt in (0, 1):
        p = oracle.Ploc(s.tris, mt, 16, 4)
        check_ploc_invariants(p, s.tris, 4)
        rays = scenes.camera_rays(s.camera, 64, 64)
        assert (p.trace(rays) == oracle.trace_brute(s.tris, rays)).all()
        # PLOC tree quality vs the reference's binned SAH tree, same cost formula
        sb = oracle.SBvh(s.tris)
        assert p.sah() <= 1.10 * sb.sah()
