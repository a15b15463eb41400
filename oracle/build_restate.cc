// Oracle — TEST INFRASTRUCTURE ONLY (see orc.h).
// Sequential restatement of the engine's BVH build: scene/centroid bounds, Morton30 and
// EMC60 codes, stable key order, PLOC clustering (Meister & Bittner 2017), SAH leaf
// collapse and the final two-child node layout. The reference has no implementation of
// any of these (src/gpu/acceleration/ is empty; README.md:154-161 lists them as planned),
// so this file is the bit-exact specification the HIP kernels are checked against.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "orc.h"

namespace orc {

static inline uint32_t FloatBits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

Tri LoadTri(const float* p) {
    return Tri{{Canon(p[0]), Canon(p[1]), Canon(p[2])},
               {Canon(p[3]), Canon(p[4]), Canon(p[5])},
               {Canon(p[6]), Canon(p[7]), Canon(p[8])}};
}

// Same expression as src/scene/triangle.h:36.
Vec3 CentroidC(const Tri& t) { return (t.v0 + t.v1 + t.v2) / 3.0f; }

static void TriBox(const Tri& t, float lo[3], float hi[3]) {
    for (int a = 0; a < 3; a++) {
        lo[a] = Min2(Min2(t.v0[a], t.v1[a]), t.v2[a]);
        hi[a] = Max2(Max2(t.v0[a], t.v1[a]), t.v2[a]);
    }
}

Bounds ComputeBounds(const float* tri_xyz, int64_t n) {
    Bounds b;
    for (int a = 0; a < 3; a++) {
        b.vmin[a] = b.cmin[a] = kInfinity;
        b.vmax[a] = b.cmax[a] = -kInfinity;
    }
    for (int64_t i = 0; i < n; i++) {
        Tri t = LoadTri(tri_xyz + 9 * i);
        float lo[3], hi[3];
        TriBox(t, lo, hi);
        Vec3 c = CentroidC(t);
        for (int a = 0; a < 3; a++) {
            b.vmin[a] = Min2(b.vmin[a], lo[a]);
            b.vmax[a] = Max2(b.vmax[a], hi[a]);
            b.cmin[a] = Min2(b.cmin[a], c[a]);
            b.cmax[a] = Max2(b.cmax[a], c[a]);
        }
    }
    return b;
}

// Spreads the low 10 bits of v so that bit k moves to bit 3k.
uint32_t ExpandBits10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Quantises (c - cmin) * scale to [0, maxq]; the input is never negative because cmin is
// the minimum of the same centroids.
static inline uint32_t Quantize(float c, float cmin, float scale, float maxq) {
    float q = (c - cmin) * scale;
    if (!(q < maxq)) q = maxq;
    return static_cast<uint32_t>(q);
}

// 30-bit Morton code: 10 bits per axis of the centroid quantised in the centroid box,
// x in the most significant position of every bit triple.
uint32_t Morton30(const Vec3& c, const Bounds& b) {
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        float ext = b.cmax[a] - b.cmin[a];
        float scale = ext > 0.0f ? 1024.0f / ext : 0.0f;
        q[a] = Quantize(c[a], b.cmin[a], scale, 1023.0f);
    }
    return (ExpandBits10(q[0]) << 2) | (ExpandBits10(q[1]) << 1) | ExpandBits10(q[2]);
}

// EMC60 bit layout (after Vinkler, Bittner, Havran 2017, "Extended Morton Codes for High
// Performance Bounding Volume Hierarchy Construction"): 48 spatial bits allocated
// adaptively — each next bit goes to the axis whose current cell is longest (ties to the
// lower axis, at most 24 bits per axis, zero-extent axes never chosen) — and 12 size bits.
// The 60-bit code is 12 groups of [1 size bit, 4 spatial bits], most significant first.
EmcLayout MakeEmcLayout(const Bounds& b) {
    EmcLayout L;
    float ext[3], cell[3];
    for (int a = 0; a < 3; a++) {
        ext[a] = b.cmax[a] - b.cmin[a];
        cell[a] = ext[a];
        L.nbits[a] = 0;
    }
    for (int t = 0; t < 48; t++) {
        int best = -1;
        for (int a = 0; a < 3; a++) {
            if (!(ext[a] > 0.0f) || L.nbits[a] >= 24) continue;
            if (best < 0 || cell[a] > cell[best]) best = a;
        }
        L.axis_of[t] = static_cast<int8_t>(best);
        if (best >= 0) {
            L.nbits[best]++;
            cell[best] = cell[best] * 0.5f;
        }
    }
    for (int a = 0; a < 3; a++) {
        L.scale[a] = ext[a] > 0.0f ? static_cast<float>(1u << L.nbits[a]) / ext[a] : 0.0f;
    }
    return L;
}

// Size measure: squared triangle-box diagonal over squared scene diagonal, mapped through
// its IEEE bit pattern (monotone, piecewise-linear in log2) to 12 bits covering 40 octaves.
static inline uint32_t SizeBits12(const float lo[3], const float hi[3], const Bounds& b) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    float d2 = dx * dx + dy * dy + dz * dz;
    float sx = b.vmax[0] - b.vmin[0], sy = b.vmax[1] - b.vmin[1], sz = b.vmax[2] - b.vmin[2];
    float s2 = sx * sx + sy * sy + sz * sz;
    float r = s2 > 0.0f ? d2 / s2 : 0.0f;
    uint32_t bits = FloatBits(r);
    const uint32_t base = 0x2B800000u;  // 2^-40
    if (bits <= base) return 0;
    uint32_t q = (bits - base) >> 17;
    return q > 4095u ? 4095u : q;
}

uint64_t Emc60(const Tri& t, const Vec3& c, const Bounds& b, const EmcLayout& L) {
    uint32_t q[3];
    int rem[3];
    for (int a = 0; a < 3; a++) {
        float maxq = static_cast<float>((1u << L.nbits[a]) - 1u);
        q[a] = L.nbits[a] > 0 ? Quantize(c[a], b.cmin[a], L.scale[a], maxq) : 0u;
        rem[a] = L.nbits[a];
    }
    float lo[3], hi[3];
    TriBox(t, lo, hi);
    uint32_t s = SizeBits12(lo, hi, b);
    uint64_t code = 0;
    for (int g = 0; g < 12; g++) {
        code = (code << 1) | ((s >> (11 - g)) & 1u);
        for (int k = 0; k < 4; k++) {
            int a = L.axis_of[4 * g + k];
            uint64_t bit = 0;
            if (a >= 0) {
                rem[a]--;
                bit = (q[a] >> rem[a]) & 1u;
            }
            code = (code << 1) | bit;
        }
    }
    return code;
}

// SA of the union of two boxes: std::min/max-form union, size = max - min, then
// 2 * (x*y + y*z + x*z) as in src/acceleration/aabb.h:19. A size that overflows is clamped
// to the largest float, so the area is never NaN (no inf * 0): nearest-neighbour keys are
// then totally ordered.
float SurfaceAreaUnion(const float* alo, const float* ahi, const float* blo, const float* bhi) {
    constexpr float kMax = 3.40282347e38f;
    float sx = Min2(Max2(ahi[0], bhi[0]) - Min2(alo[0], blo[0]), kMax);
    float sy = Min2(Max2(ahi[1], bhi[1]) - Min2(alo[1], blo[1]), kMax);
    float sz = Min2(Max2(ahi[2], bhi[2]) - Min2(alo[2], blo[2]), kMax);
    return 2.0f * (sx * sy + sy * sz + sx * sz);
}

static inline float SurfaceAreaBox(const float* lo, const float* hi) {
    float sx = hi[0] - lo[0], sy = hi[1] - lo[1], sz = hi[2] - lo[2];
    return 2.0f * (sx * sy + sy * sz + sx * sz);
}

// Priority of an unordered cluster pair (same value for (i,j) and (j,i)).
uint32_t PairHash(int32_t i, int32_t j) {
    uint32_t lo = uint32_t(std::min(i, j)), hi = uint32_t(std::max(i, j));
    return HashU32(lo * 0x9E3779B1u ^ hi);
}

// PLOC nearest neighbour: argmin over 0 < |j - i| <= radius of the triple
// (SA(box_i U box_j), PairHash(i, j), j). Exact distance ties go to the pair of lower
// hash priority: the order is symmetric in the pair, so the globally smallest pair is
// always mutual (at least one merge per iteration), and on runs of identical boxes a
// constant fraction of clusters still merges per iteration (a lower-index rule would
// merge a single pair per iteration there).
void PlocNearestNeighbours(const std::vector<PNode>& cl, int c, int radius, std::vector<int32_t>& nn) {
    nn.assign(c, -1);
    auto work = [&](int begin, int end) {
        for (int i = begin; i < end; i++) {
            float best_d = 0.0f;
            int32_t best_j = -1;
            uint32_t best_h = 0;
            bool hv = false;
            int lo = std::max(0, i - radius), hi = std::min(c - 1, i + radius);
            for (int j = lo; j <= hi; j++) {
                if (j == i) continue;
                float d = SurfaceAreaUnion(cl[i].lo, cl[i].hi, cl[j].lo, cl[j].hi);
                if (best_j < 0 || d < best_d) {
                    best_d = d;
                    best_j = j;
                    hv = false;
                } else if (d == best_d) {
                    if (!hv) {
                        best_h = PairHash(i, best_j);
                        hv = true;
                    }
                    uint32_t h = PairHash(i, j);
                    if (h < best_h || (h == best_h && j < best_j)) {
                        best_j = j;
                        best_h = h;
                    }
                }
            }
            nn[i] = best_j;
        }
    };
    int nt = c > 65536 ? 8 : 1;
    if (nt == 1) {
        work(0, c);
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; t++) th.emplace_back(work, int(int64_t(c) * t / nt), int(int64_t(c) * (t + 1) / nt));
        for (auto& x : th) x.join();
    }
}

static void BuildFinalTree(const float* tri_xyz, int64_t n, const BuildParams& p, PlocResult& R) {
    const int64_t nn = 2 * n - 1;
    const int root = int(nn - 1);
    const int maxleaf = p.max_leaf;
    FinalTree& T = R.tree;
    for (int a = 0; a < 3; a++) {
        T.root_lo[a] = R.nodes[root].lo[a];
        T.root_hi[a] = R.nodes[root].hi[a];
    }
    T.nodes.clear();
    T.prim_order.assign(n, -1);
    T.tri_ordered.assign(12 * n, 0.0f);

    // minimum sorted leaf position below each node: orders the final leaves
    std::vector<int32_t> minleaf(nn);
    for (int64_t k = 0; k < n; k++) minleaf[k] = int32_t(k);
    for (int64_t id = n; id < nn; id++) minleaf[id] = std::min(minleaf[R.nodes[id].left], minleaf[R.nodes[id].right]);

    std::vector<uint8_t> inside(nn, 0);
    for (int64_t id = 0; id < nn; id++) {
        int q = R.parent[id];
        while (q >= 0 && R.count[q] <= maxleaf) {
            if (R.collapsed[q]) {
                inside[id] = 1;
                break;
            }
            q = R.parent[q];
        }
    }
    auto is_final_leaf = [&](int64_t id) { return R.collapsed[id] && !inside[id]; };
    auto is_kept_inner = [&](int64_t id) { return id >= n && !R.collapsed[id] && !inside[id]; };

    // prim ranges: final leaves in order of their minimum sorted position
    std::vector<int32_t> slot(n, 0);  // count of the final leaf whose minleaf is k
    for (int64_t id = 0; id < nn; id++)
        if (is_final_leaf(id)) slot[minleaf[id]] = R.count[id];
    std::vector<int32_t> begin_at(n, 0);
    int32_t acc = 0;
    for (int64_t k = 0; k < n; k++) {
        begin_at[k] = acc;
        acc += slot[k];
    }
    // gather prims of each final leaf: left-first DFS of its PLOC subtree
    for (int64_t id = 0; id < nn; id++) {
        if (!is_final_leaf(id)) continue;
        int32_t b = begin_at[minleaf[id]];
        std::vector<int32_t> st{int32_t(id)};
        while (!st.empty()) {
            int32_t x = st.back();
            st.pop_back();
            if (x < n) {
                T.prim_order[b++] = R.nodes[x].left;
            } else {
                st.push_back(R.nodes[x].right);
                st.push_back(R.nodes[x].left);
            }
        }
    }
    for (int64_t k = 0; k < n; k++) {
        int32_t orig = T.prim_order[k];
        Tri t = LoadTri(tri_xyz + 9 * int64_t(orig));
        float* d = &T.tri_ordered[12 * k];
        d[0] = t.v0.x, d[1] = t.v0.y, d[2] = t.v0.z;
        std::memcpy(&d[3], &orig, 4);
        d[4] = t.v1.x, d[5] = t.v1.y, d[6] = t.v1.z, d[7] = 0.0f;
        d[8] = t.v2.x, d[9] = t.v2.y, d[10] = t.v2.z, d[11] = 0.0f;
    }
    auto leaf_ref = [&](int64_t id) { return ~((begin_at[minleaf[id]] << 4) | R.count[id]); };

    if (R.collapsed[root]) {  // the whole scene is one leaf: one node, right slot empty
        FinalTree::Node nd{};
        const PNode& r = R.nodes[root];
        nd.lx0 = nd.rx0 = r.lo[0];
        nd.lx1 = nd.rx1 = r.hi[0];
        nd.ly0 = nd.ry0 = r.lo[1];
        nd.ly1 = nd.ry1 = r.hi[1];
        nd.lz0 = nd.rz0 = r.lo[2];
        nd.lz1 = nd.rz1 = r.hi[2];
        nd.left = leaf_ref(root);
        nd.right = ~0;  // leaf with zero primitives
        T.nodes.push_back(nd);
        return;
    }
    // kept inner nodes, numbered by descending PLOC id (root = 0)
    std::vector<int32_t> idx(nn, -1);
    int32_t K = 0;
    for (int64_t id = n; id < nn; id++)
        if (is_kept_inner(id)) K++;
    int32_t rank = 0;
    for (int64_t id = n; id < nn; id++)
        if (is_kept_inner(id)) idx[id] = K - 1 - rank++;
    T.nodes.assign(K, FinalTree::Node{});
    for (int64_t id = n; id < nn; id++) {
        if (!is_kept_inner(id)) continue;
        FinalTree::Node& nd = T.nodes[idx[id]];
        const PNode& L = R.nodes[R.nodes[id].left];
        const PNode& Rn = R.nodes[R.nodes[id].right];
        nd.lx0 = L.lo[0], nd.lx1 = L.hi[0], nd.ly0 = L.lo[1], nd.ly1 = L.hi[1];
        nd.rx0 = Rn.lo[0], nd.rx1 = Rn.hi[0], nd.ry0 = Rn.lo[1], nd.ry1 = Rn.hi[1];
        nd.lz0 = L.lo[2], nd.lz1 = L.hi[2], nd.rz0 = Rn.lo[2], nd.rz1 = Rn.hi[2];
        int32_t l = R.nodes[id].left, r = R.nodes[id].right;
        nd.left = is_kept_inner(l) ? idx[l] : leaf_ref(l);
        nd.right = is_kept_inner(r) ? idx[r] : leaf_ref(r);
        nd.pad0 = nd.pad1 = 0;
    }
}

void PlocBuild(const float* tri_xyz, int64_t n, const BuildParams& p, PlocResult& R) {
    R = PlocResult{};
    R.bounds = ComputeBounds(tri_xyz, n);
    const Bounds& B = R.bounds;
    R.codes.resize(n);
    EmcLayout L = MakeEmcLayout(B);
    for (int64_t i = 0; i < n; i++) {
        Tri t = LoadTri(tri_xyz + 9 * i);
        Vec3 c = CentroidC(t);
        R.codes[i] = p.morton_type == 0 ? uint64_t(Morton30(c, B)) : Emc60(t, c, B, L);
    }
    // stable order by key (equal keys keep input order)
    R.sorted_idx.resize(n);
    std::iota(R.sorted_idx.begin(), R.sorted_idx.end(), 0u);
    std::stable_sort(R.sorted_idx.begin(), R.sorted_idx.end(),
                     [&](uint32_t a, uint32_t b) { return R.codes[a] < R.codes[b]; });
    R.sorted_codes.resize(n);
    for (int64_t k = 0; k < n; k++) R.sorted_codes[k] = R.codes[R.sorted_idx[k]];

    const int64_t nn = 2 * n - 1;
    R.nodes.assign(nn, PNode{});
    R.parent.assign(nn, -1);
    std::vector<PNode> cl(n);  // cluster list: box + node id in `left`
    for (int64_t k = 0; k < n; k++) {
        Tri t = LoadTri(tri_xyz + 9 * int64_t(R.sorted_idx[k]));
        PNode& nd = R.nodes[k];
        TriBox(t, nd.lo, nd.hi);
        nd.left = int32_t(R.sorted_idx[k]);
        nd.right = -1;
        cl[k] = nd;
        cl[k].left = int32_t(k);
    }
    int c = int(n);
    std::vector<int32_t> nnb;
    std::vector<PNode> next;
    next.reserve(n);
    while (c > 1) {
        R.iter_counts.push_back(c);
        PlocNearestNeighbours(cl, c, p.radius, nnb);
        next.clear();
        int32_t base = int32_t(2 * n - c), merges = 0;
        for (int i = 0; i < c; i++) {
            int j = nnb[i];
            if (nnb[j] == i) {
                if (i < j) {
                    int32_t id = base + merges++;
                    PNode& nd = R.nodes[id];
                    for (int a = 0; a < 3; a++) {
                        nd.lo[a] = Min2(cl[i].lo[a], cl[j].lo[a]);
                        nd.hi[a] = Max2(cl[i].hi[a], cl[j].hi[a]);
                    }
                    nd.left = cl[i].left;
                    nd.right = cl[j].left;
                    R.parent[nd.left] = id;
                    R.parent[nd.right] = id;
                    PNode m = nd;
                    m.left = id;
                    next.push_back(m);
                }
            } else {
                next.push_back(cl[i]);
            }
        }
        cl.swap(next);
        c = int(cl.size());
    }

    // bottom-up SAH collapse (Ct = Ci = 1): become a leaf when it holds at most max_leaf
    // triangles and SA * count <= SA + cost(left) + cost(right)
    R.cost.assign(nn, 0.0f);
    R.count.assign(nn, 0);
    R.collapsed.assign(nn, 0);
    for (int64_t k = 0; k < n; k++) {
        R.count[k] = 1;
        R.cost[k] = SurfaceAreaBox(R.nodes[k].lo, R.nodes[k].hi) * 1.0f;
        R.collapsed[k] = 1;
    }
    for (int64_t id = n; id < nn; id++) {
        const PNode& nd = R.nodes[id];
        int32_t cnt = R.count[nd.left] + R.count[nd.right];
        float sa = SurfaceAreaBox(nd.lo, nd.hi);
        float inner = sa + R.cost[nd.left] + R.cost[nd.right];
        float cost = inner;
        uint8_t col = 0;
        if (cnt <= p.max_leaf) {
            float leafc = sa * float(cnt);
            if (leafc <= inner) {
                cost = leafc;
                col = 1;
            }
        }
        R.count[id] = cnt;
        R.cost[id] = cost;
        R.collapsed[id] = col;
    }
    BuildFinalTree(tri_xyz, n, p, R);
    if (p.bv_type == 1) {
        const SobbSet S = MakeSobbSet(p.kdop_size);
        const float pad = KdopPad(R.bounds);
        const size_t w = size_t(2 * S.ndirs);
        R.kdop.assign(size_t(nn) * w, 0.0f);
        for (int64_t k = 0; k < n; k++)
            LeafKdop(S, LoadTri(tri_xyz + 9 * int64_t(R.sorted_idx[k])), pad, &R.kdop[size_t(k) * w]);
        for (int64_t id = n; id < nn; id++) {
            const float* l = &R.kdop[size_t(R.nodes[id].left) * w];
            const float* r = &R.kdop[size_t(R.nodes[id].right) * w];
            float* o = &R.kdop[size_t(id) * w];
            for (int k = 0; k < S.ndirs; k++) {
                o[k] = Min2(l[k], r[k]);
                o[S.ndirs + k] = Max2(l[S.ndirs + k], r[S.ndirs + k]);
            }
        }
        BuildSobbNodes(S, R, R.kdop, n, p.max_leaf, R.tree.nodes);
        R.tree.sobb = true;
        R.tree.set = S;
    }
}

// SOBB surface area 2 (wa wb + wb wc + wa wc) / |det| in FP64 for child s (0 left, 1 right).
static double SobbArea(const FinalTree& t, const FinalTree::Node& nd, int s) {
    const float* f = &nd.lx0 + 6 * s;
    const uint32_t d = (uint32_t(nd.pad0) >> (12 * s)) & 0xFFFu;
    const int a = d & 15, b = (d >> 4) & 15, c = (d >> 8) & 15;
    const float* n1 = t.set.dirs[a];
    const float* n2 = t.set.dirs[b];
    const float* n3 = t.set.dirs[c];
    const double det = double(n1[0]) * (double(n2[1]) * n3[2] - double(n2[2]) * n3[1]) +
                       double(n1[1]) * (double(n2[2]) * n3[0] - double(n2[0]) * n3[2]) +
                       double(n1[2]) * (double(n2[0]) * n3[1] - double(n2[1]) * n3[0]);
    const double wa = double(f[1]) - f[0], wb = double(f[3]) - f[2], wc = double(f[5]) - f[4];
    return 2.0 * (wa * wb + wb * wc + wa * wc) / std::fabs(det);
}

static double AabbArea(const FinalTree::Node& nd, int s) {
    const double x0 = s ? nd.rx0 : nd.lx0, x1 = s ? nd.rx1 : nd.lx1;
    const double y0 = s ? nd.ry0 : nd.ly0, y1 = s ? nd.ry1 : nd.ly1;
    const double z0 = s ? nd.rz0 : nd.lz0, z1 = s ? nd.rz1 : nd.lz1;
    const double sx = x1 - x0, sy = y1 - y0, sz = z1 - z0;
    return 2.0 * (sx * sy + sy * sz + sx * sz);
}

// Ct * SA(root) + sum over the children of every node of (Ct SA for an inner child,
// Ci SA |leaf| for a leaf), divided by SA(root); SA is the child's AABB or SOBB area and
// SA(root) the area of the root's axis-aligned box (both trees share the scene bounds).
double SahCostFinal(const FinalTree& t, double ct, double ci) {
    if (t.nodes.empty()) return 0.0;
    const double rx = double(t.root_hi[0]) - t.root_lo[0], ry = double(t.root_hi[1]) - t.root_lo[1],
                 rz = double(t.root_hi[2]) - t.root_lo[2];
    const double root = 2.0 * (rx * ry + ry * rz + rx * rz);
    if (!(root > 0)) return 0.0;
    auto area = [&](const FinalTree::Node& nd, int s) { return t.sobb ? SobbArea(t, nd, s) : AabbArea(nd, s); };
    const FinalTree::Node& r = t.nodes[0];
    if (r.right == ~0 && r.left < 0) return ci * ((~r.left) & 15);  // one leaf: Ci * |leaf| * SA / SA
    double sum = ct * root;
    for (const auto& nd : t.nodes) {
        for (int s = 0; s < 2; s++) {
            const int32_t ref = s == 0 ? nd.left : nd.right;
            const double a = area(nd, s);
            sum += ref >= 0 ? ct * a : ci * a * ((~ref) & 15);
        }
    }
    return sum / root;
}

}  // namespace orc

