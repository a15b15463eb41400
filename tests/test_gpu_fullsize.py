"""GPU parity at the full sizes of the 2M-triangle mixed scene and of the soups large enough
to run every multi-workgroup PLOC iteration instance in both its one- and two-cluster-per-thread
forms (block 128, 256 and 512), bit for bit against the oracle: codes, sorted order, every
iteration's cluster count, the PLOC node array (merge list), the final nodes, the triangle
order and the SAH. A million of the scene's 4,194,304 random rays are traced over the full tree
and compared with the oracle's traversal of the same tree and with an FP64 intersection."""
import sys

import numpy as np
import pytest

from diplodocus_amd import scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def soup(n, seed, spread=10.0, size=0.02):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, spread, (n, 1, 3))
    return (c + rng.normal(0, size, (n, 3, 3))).reshape(n, 9).astype(np.float32)


@pytest.fixture(scope="module")
def mixed():
    s = scenes.mixed_2m()
    assert s.n == 2_000_002
    return s


def build_and_compare(oracle, ctx, tris, mt, bv, radius, block=0, digit=0, keep_tree=False):
    from diplodocus_amd import accel

    say(f"full-size case n={len(tris)} mt={mt} bv={bv} r={radius} block={block} digit={digit}")
    p = accel.make_params(morton_type=mt, radius=radius, max_leaf=4, block_size=block, sort_digit_bits=digit,
                          keep_intermediates=True, bv_type=bv, kdop_size=32)
    g = accel.GpuBvh(ctx).build(np.ascontiguousarray(tris, np.float32), p)
    d = g.download(intermediates=True)
    o = oracle.Ploc(tris, mt, radius, 4, bv_type=bv, kdop_size=32)
    assert (d["codes"] == o.codes).all(), "codes"
    assert (d["sorted_idx"] == o.sorted_idx).all(), "order"
    assert len(d["iter_counts"]) == len(o.iter_counts), "iterations"
    assert (d["iter_counts"] == o.iter_counts).all(), "iteration counts"
    pn = d["ploc_nodes"]
    assert (pn["lo"].view(np.uint32) == o.nodes["lo"].view(np.uint32)).all(), "node boxes"
    assert (pn["hi"].view(np.uint32) == o.nodes["hi"].view(np.uint32)).all(), "node boxes"
    assert (pn["left"] == o.nodes["left"]).all() and (pn["right"] == o.nodes["right"]).all(), "merge list"
    assert len(d["nodes"]) == len(o.fnodes)
    assert (d["nodes"].view(np.uint32) == o.fnodes.view(np.uint32)).all(), "final nodes"
    assert (d["prim_order"] == o.prim_order).all(), "prim order"
    assert (d["tri_ordered"].view(np.uint32) == o.tri_ordered.view(np.uint32)).all(), "ordered triangles"
    assert g.stats.ploc_iterations == len(o.iter_counts)
    assert g.stats.sah_cost == pytest.approx(o.sah(), rel=1e-9)
    say(f"  equal: {len(o.iter_counts)} iterations, {len(o.fnodes)} nodes, SAH {o.sah():.4f}")
    if keep_tree:
        return g, d
    g.close()
    return None, None


@pytest.mark.parametrize("mt,bv,radius", [(1, 1, 16), (0, 0, 8), (0, 0, 32), (1, 0, 16), (0, 1, 16)])
def test_mixed2m_full_build_bitexact(oracle, ctx, mixed, mt, bv, radius):
    """The headline build (PLOC+EMC+SOBB r16) and the other variants at the full 2,000,002
    triangles: every PLOC iteration kernel instance the headline build launches (the two-launch
    iterations above ~98k clusters, the fused four-lane iterations below, the tail) runs here."""
    build_and_compare(oracle, ctx, mixed.tris, mt, bv, radius)


@pytest.mark.parametrize("block,digit", [(128, 8), (512, 4), (256, 4)])
def test_large_soup_all_blocks_bitexact(oracle, ctx, block, digit):
    """1.2M triangles: blocks of 128 and 512 threads take two clusters per thread from 1,048,576
    clusters on a 256-CU device (4096 and 1024 workgroups), 256 from 524,288."""
    build_and_compare(oracle, ctx, soup(1_200_000, block + digit), 0, 0, 16, block=block, digit=digit)


@pytest.mark.parametrize("radius,mt,bv", [(1, 1, 1), (32, 1, 1), (2, 0, 0), (4, 0, 1), (5, 1, 0)])
def test_two_launch_iterations_radius_extremes(oracle, ctx, radius, mt, bv):
    """300k triangles: the iterations above ~98k clusters run as a neighbour launch and a merge
    launch; at the smallest and largest radii the tiles' halos (2 r on each side) and the range
    of searched clusters reach their bounds. From radius 4 up (and at most 32 KiB of areas)
    the neighbour launch computes each pair's area once and shares it between the pair's two
    clusters: radius 4 is the smallest such radius, radius 5 leaves a partial group of four
    candidates at the end of each row."""
    build_and_compare(oracle, ctx, soup(300_000, 7 + radius), mt, bv, radius)


def test_mixed2m_full_trace_matches_oracle(oracle, ctx, mixed):
    """A million of the bench's random rays over the full headline tree: every hit equals the
    oracle's ordered traversal of the same tree bit for bit, and t agrees with an FP64
    intersection of the returned triangle to a relative 1e-5."""
    g, d = build_and_compare(oracle, ctx, mixed.tris, 1, 1, 16, keep_tree=True)
    rays = scenes.random_rays_towards(mixed.tris, 4_194_304, seed=11)[:1_048_576]
    say("trace 1,048,576 rays over the full tree")
    h = g.trace(rays)
    ref = oracle.trace_final(d["nodes"], d["tri_ordered"], rays, 0, sobb_kdop=32)
    eq = (h.view(np.uint32).reshape(-1, 4) == ref.view(np.uint32).reshape(-1, 4)).all(axis=1)
    assert eq.all(), f"{int((~eq).sum())} hits differ"
    hit = np.nonzero(h["prim"] >= 0)[0]
    assert len(hit) > len(rays) // 2
    sel = hit[:: max(1, len(hit) // 20000)]
    t64 = np.array([oracle.intersect_fp64(mixed.tris[h["prim"][k]], rays[k])[0] for k in sel])
    rel = np.abs(h["t"][sel].astype(np.float64) - t64) / np.maximum(t64, 1e-30)
    assert rel.max() <= 1e-5, rel.max()
    g.close()
