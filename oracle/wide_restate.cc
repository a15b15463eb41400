// Oracle — TEST INFRASTRUCTURE ONLY (see orc.h).
//
// Wide trees (4 or 8 children per node) with quantized child boxes, made from the two-child
// final tree, and their traversal. The reference has no wide tree: its GPU path stops at a
// brute-force intersector (src/gpu/renderer/gpu_renderer_util.h:29-94) and its CPU SBvh is
// two-child (src/acceleration/sbvh.h:25-31). Results must equal the two-child trace's
// (closest hit: smallest t, ties to the lower primitive id), which is how the restatement is
// pinned, together with the containment, tightness and greedy-rule checks in tests/.
//
//   RefitAabb  axis-aligned child boxes of every node of a final tree, from its triangles
//              (for SOBB trees; an AABB tree's own boxes are reproduced exactly)
//   Widen      wide node i = node i's subtree opened greedily (largest child surface area
//              first, ties to the leftmost) until it has W children; children keep their
//              left-to-right order; each box is stored as 8-bit offsets from the node's
//              minimum corner in units of a power of two per axis
//   TraceWide  ordered traversal: hit children visited by (tnear, slot)
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "orc.h"

namespace orc {

namespace {

bool EmptyLeaf(int32_t ref) { return ref < 0 && ((~ref) & 15) == 0; }

void LeafBox(const FinalTree& t, int32_t ref, float lo[3], float hi[3]) {
    const int32_t packed = ~ref, b0 = packed >> 4, cnt = packed & 15;
    for (int a = 0; a < 3; a++) lo[a] = kInfinity, hi[a] = -kInfinity;
    for (int32_t k = b0; k < b0 + cnt; k++) {
        const float* d = &t.tri_ordered[12 * size_t(k)];
        for (int v = 0; v < 3; v++)
            for (int a = 0; a < 3; a++) {
                lo[a] = Min2(lo[a], d[4 * v + a]);
                hi[a] = Max2(hi[a], d[4 * v + a]);
            }
    }
}

void SetChild(FinalTree::Node& n, int side, const float lo[3], const float hi[3]) {
    if (side == 0) {
        n.lx0 = lo[0], n.lx1 = hi[0], n.ly0 = lo[1], n.ly1 = hi[1], n.lz0 = lo[2], n.lz1 = hi[2];
    } else {
        n.rx0 = lo[0], n.rx1 = hi[0], n.ry0 = lo[1], n.ry1 = hi[1], n.rz0 = lo[2], n.rz1 = hi[2];
    }
}

void GetChild(const FinalTree::Node& n, int side, float lo[3], float hi[3]) {
    if (side == 0) {
        lo[0] = n.lx0, hi[0] = n.lx1, lo[1] = n.ly0, hi[1] = n.ly1, lo[2] = n.lz0, hi[2] = n.lz1;
    } else {
        lo[0] = n.rx0, hi[0] = n.rx1, lo[1] = n.ry0, hi[1] = n.ry1, lo[2] = n.rz0, hi[2] = n.rz1;
    }
}

// Box of the subtree below `ref`; fills the child boxes of every inner node on the way.
void RefitRec(const FinalTree& t, int32_t ref, std::vector<FinalTree::Node>& out, float lo[3], float hi[3],
              int depth) {
    if (depth > 4096) throw std::runtime_error("RefitAabb: tree too deep");
    if (ref < 0) {
        LeafBox(t, ref, lo, hi);
        return;
    }
    const FinalTree::Node& nd = t.nodes[size_t(ref)];
    FinalTree::Node& o = out[size_t(ref)];
    o.left = nd.left, o.right = nd.right, o.pad0 = 0, o.pad1 = 0;
    float llo[3], lhi[3], rlo[3], rhi[3];
    RefitRec(t, nd.left, out, llo, lhi, depth + 1);
    if (EmptyLeaf(nd.right)) {  // a one-leaf tree: the empty right slot repeats the left box
        std::memcpy(rlo, llo, sizeof llo), std::memcpy(rhi, lhi, sizeof lhi);
    } else {
        RefitRec(t, nd.right, out, rlo, rhi, depth + 1);
    }
    SetChild(o, 0, llo, lhi);
    SetChild(o, 1, rlo, rhi);
    for (int a = 0; a < 3; a++) lo[a] = Min2(llo[a], rlo[a]), hi[a] = Max2(lhi[a], rhi[a]);
}

struct Ent {
    int32_t ref;
    float lo[3], hi[3];
};

// Half the surface area, FP64 from the FP32 bounds.
double HalfArea(const Ent& e) {
    const double dx = double(e.hi[0]) - double(e.lo[0]);
    const double dy = double(e.hi[1]) - double(e.lo[1]);
    const double dz = double(e.hi[2]) - double(e.lo[2]);
    return dx * dy + dy * dz + dz * dx;
}

float ScaleOf(uint32_t b) {
    const uint32_t u = b << 23;
    float s;
    std::memcpy(&s, &u, 4);
    return s;
}

// Smallest biased exponent b in [1, 254] with fmaf(255, 2^(b-127), L) >= H.
uint32_t ChooseExponent(float L, float H) {
    const double ext = double(H) - double(L);
    int b = 1;
    if (ext > 0.0) {
        int e = 0;
        std::frexp(ext / 255.0, &e);
        b = std::min(254, std::max(1, e + 127));
    }
    while (b > 1 && std::fmaf(255.0f, ScaleOf(uint32_t(b - 1)), L) >= H) b--;
    while (b < 254 && !(std::fmaf(255.0f, ScaleOf(uint32_t(b)), L) >= H)) b++;
    return uint32_t(b);
}

// Largest q in [0, 255] with fmaf(q, s, L) <= lo.
uint32_t QuantLo(float lo, float L, float s) {
    double est = std::floor((double(lo) - double(L)) / double(s));
    int q = int(std::min(255.0, std::max(0.0, est)));
    while (q < 255 && std::fmaf(float(q + 1), s, L) <= lo) q++;
    while (q > 0 && std::fmaf(float(q), s, L) > lo) q--;
    return uint32_t(q);
}

// Smallest q in [0, 255] with fmaf(q, s, L) >= hi.
uint32_t QuantHi(float hi, float L, float s) {
    double est = std::ceil((double(hi) - double(L)) / double(s));
    int q = int(std::min(255.0, std::max(0.0, est)));
    while (q > 0 && std::fmaf(float(q - 1), s, L) >= hi) q--;
    while (q < 255 && !(std::fmaf(float(q), s, L) >= hi)) q++;
    return uint32_t(q);
}

}  // namespace

int WideNodeBytes(int width) { return width == 4 ? 64 : 96; }

void RefitAabb(const FinalTree& t, std::vector<FinalTree::Node>& out) {
    out.assign(t.nodes.size(), FinalTree::Node{});
    if (t.nodes.empty()) return;
    float lo[3], hi[3];
    RefitRec(t, 0, out, lo, hi, 0);
}

void Widen(const std::vector<FinalTree::Node>& nodes, int width, std::vector<uint8_t>& out) {
    if (width != 4 && width != 8) throw std::runtime_error("Widen: width must be 4 or 8");
    const int W = width, NB = WideNodeBytes(W);
    out.assign(nodes.size() * size_t(NB), 0);
    for (size_t i = 0; i < nodes.size(); i++) {
        Ent fr[8];
        int c = 0;
        auto child = [&](const FinalTree::Node& nd, int side) {
            Ent e;
            e.ref = side == 0 ? nd.left : nd.right;
            GetChild(nd, side, e.lo, e.hi);
            return e;
        };
        for (int side = 0; side < 2; side++) {
            const Ent e = child(nodes[i], side);
            if (!EmptyLeaf(e.ref)) fr[c++] = e;
        }
        while (c < W) {
            int best = -1;
            double ba = 0.0;
            for (int p = 0; p < c; p++) {
                if (fr[p].ref < 0 || size_t(fr[p].ref) >= nodes.size()) continue;
                const double a = HalfArea(fr[p]);
                if (best < 0 || a > ba) best = p, ba = a;
            }
            if (best < 0) break;
            const FinalTree::Node& g = nodes[size_t(fr[best].ref)];
            for (int p = c; p > best + 1; p--) fr[p] = fr[p - 1];
            fr[best] = child(g, 0);
            fr[best + 1] = child(g, 1);
            c++;
        }
        uint8_t* o = &out[i * size_t(NB)];
        float origin[3];
        uint32_t bexp[3];
        float s[3];
        for (int a = 0; a < 3; a++) {
            float L = fr[0].lo[a], H = fr[0].hi[a];
            for (int p = 1; p < c; p++) L = Min2(L, fr[p].lo[a]), H = Max2(H, fr[p].hi[a]);
            origin[a] = c > 0 ? L : 0.0f;
            bexp[a] = c > 0 ? ChooseExponent(L, H) : 1u;
            s[a] = ScaleOf(bexp[a]);
        }
        std::memcpy(o, origin, 12);
        const uint32_t meta = bexp[0] | (bexp[1] << 8) | (bexp[2] << 16) | (uint32_t(c) << 24);
        std::memcpy(o + 12, &meta, 4);
        for (int k = 0; k < W; k++) {
            const int32_t r = k < c ? fr[k].ref : ~0;
            std::memcpy(o + 16 + 4 * k, &r, 4);
        }
        uint8_t* q = o + 16 + 4 * W;  // planes: lo x, hi x, lo y, hi y, lo z, hi z; byte k = slot k
        for (int k = 0; k < c; k++)
            for (int a = 0; a < 3; a++) {
                q[(2 * a) * W + k] = uint8_t(QuantLo(fr[k].lo[a], origin[a], s[a]));
                q[(2 * a + 1) * W + k] = uint8_t(QuantHi(fr[k].hi[a], origin[a], s[a]));
            }
    }
}

void DecodeWideChild(const uint8_t* node, int width, int k, float lo[3], float hi[3]) {
    float origin[3];
    uint32_t meta;
    std::memcpy(origin, node, 12);
    std::memcpy(&meta, node + 12, 4);
    const uint8_t* q = node + 16 + 4 * width;
    for (int a = 0; a < 3; a++) {
        const float s = ScaleOf((meta >> (8 * a)) & 255u);
        lo[a] = std::fmaf(float(q[(2 * a) * width + k]), s, origin[a]);
        hi[a] = std::fmaf(float(q[(2 * a + 1) * width + k]), s, origin[a]);
    }
}

void TraceWideOne(const WideTree& T, const RayRec& r, uint32_t flags, HitRec& h, TraceCount* count) {
    const bool any = flags & 1u, cull = flags & 2u;
    const int W = T.width, NB = WideNodeBytes(W);
    const WtRay w = PrepareRay(r);
    h = HitRec{0.0f, 0.0f, 0.0f, -1};
    if (T.node_count <= 0) return;
    float best_t = r.tmax;
    std::vector<int32_t> stack;
    stack.push_back(0);
    const int64_t nprim = int64_t(T.tri_ordered.size() / 12);
    while (!stack.empty()) {
        const int32_t ref = stack.back();
        stack.pop_back();
        if (ref >= 0) {
            if (ref >= T.node_count) break;
            if (count) count->nodes++;
            const uint8_t* nd = &T.nodes[size_t(ref) * size_t(NB)];
            uint32_t meta;
            std::memcpy(&meta, nd + 12, 4);
            const int c = int(meta >> 24);
            struct Hit {
                float t;
                int k;
                int32_t ref;
            } hs[8];
            int nh = 0;
            for (int k = 0; k < c && k < W; k++) {
                float lo[3], hi[3], tn;
                DecodeWideChild(nd, W, k, lo, hi);
                if (!IntersectBoxRobust(w, r.tmin, best_t, lo, hi, tn)) continue;
                int32_t cr;
                std::memcpy(&cr, nd + 16 + 4 * k, 4);
                hs[nh++] = Hit{tn, k, cr};
            }
            std::sort(hs, hs + nh, [](const Hit& a, const Hit& b) { return a.t < b.t || (a.t == b.t && a.k < b.k); });
            for (int j = nh - 1; j >= 0; j--) stack.push_back(hs[j].ref);
        } else {
            const int32_t packed = ~ref, b0 = packed >> 4;
            const int32_t cnt = int64_t(b0) + (packed & 15) <= nprim ? (packed & 15) : 0;
            for (int32_t k = b0; k < b0 + cnt; k++) {
                const float* d = &T.tri_ordered[12 * size_t(k)];
                int32_t prim;
                std::memcpy(&prim, &d[3], 4);
                float t, b1, b2;
                if (count) count->tris++;
                if (!IntersectWatertight(w, d, d + 4, d + 8, cull, t, b1, b2)) continue;
                if (!(t > r.tmin)) continue;
                if (any) {
                    if (t < r.tmax) {
                        h = HitRec{t, b1, b2, prim};
                        return;
                    }
                    continue;
                }
                if (t < best_t || (t == best_t && h.prim >= 0 && prim < h.prim)) {
                    best_t = t;
                    h = HitRec{t, b1, b2, prim};
                }
            }
        }
    }
}

}  // namespace orc
