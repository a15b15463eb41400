// Device helpers shared by the single-scene build kernels (build.hip) and the batched
// per-mesh build (batch.hip): triangle loads, bounds decoding, Morton30 / EMC60 codes, the
// fused leaf-collapse record, SOBB k-DOPs and fits, the PLOC nearest-neighbour search and the
// final-layout node classification. Every function follows the oracle's restatement
// operation by operation (no FMA contraction, IEEE denormals, Min2/Max2 select semantics).
#pragma once

#include <hip/hip_runtime.h>

#include "dpl_internal.h"

namespace dpl {

__device__ __forceinline__ void LoadTri(const float* tri, int64_t i, float v[9]) {
    const float* p = tri + 9 * i;
#pragma unroll
    for (int k = 0; k < 9; k++) v[k] = Canon(p[k]);
}

__device__ __forceinline__ void TriBox(const float v[9], float lo[3], float hi[3]) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lo[a] = Min2(Min2(v[a], v[3 + a]), v[6 + a]);
        hi[a] = Max2(Max2(v[a], v[3 + a]), v[6 + a]);
    }
}

// ((v0 + v1) + v2) / 3.0f per component (src/scene/triangle.h:36)
__device__ __forceinline__ void Centroid(const float v[9], float c[3]) {
#pragma unroll
    for (int a = 0; a < 3; a++) c[a] = ((v[a] + v[3 + a]) + v[6 + a]) / 3.0f;
}

__device__ __forceinline__ uint32_t ExpandBits10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t Quantize(float c, float cmin, float scale, float maxq) {
    float q = (c - cmin) * scale;
    if (!(q < maxq)) q = maxq;
    return uint32_t(q);
}

struct DecodedBounds {
    float vmin[3], vmax[3], cmin[3], cmax[3];
};

__device__ __forceinline__ DecodedBounds Decode(const uint32_t* b) {
    DecodedBounds d;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        d.vmin[a] = OrderedToFloat(b[a]);
        d.vmax[a] = OrderedToFloat(b[3 + a]);
        d.cmin[a] = OrderedToFloat(b[6 + a]);
        d.cmax[a] = OrderedToFloat(b[9 + a]);
    }
    return d;
}

// ---------------------------------------------------------------------------------------
// codes
// ---------------------------------------------------------------------------------------
struct EmcLayoutDev {
    int nbits[3];
    int8_t axis_of[48];
    float scale[3];
};

// Same allocation as the oracle's MakeEmcLayout: each next spatial bit goes to the axis
// with the longest current cell (ties to the lower axis, at most 24 bits, zero-extent
// axes never chosen).
__device__ inline void MakeEmcLayout(const DecodedBounds& d, EmcLayoutDev& L) {
    float ext[3], cell[3];
    for (int a = 0; a < 3; a++) {
        ext[a] = d.cmax[a] - d.cmin[a];
        cell[a] = ext[a];
        L.nbits[a] = 0;
    }
    for (int t = 0; t < 48; t++) {
        int best = -1;
        for (int a = 0; a < 3; a++) {
            if (!(ext[a] > 0.0f) || L.nbits[a] >= 24) continue;
            if (best < 0 || cell[a] > cell[best]) best = a;
        }
        L.axis_of[t] = int8_t(best);
        if (best >= 0) {
            L.nbits[best]++;
            cell[best] = cell[best] * 0.5f;
        }
    }
    for (int a = 0; a < 3; a++) L.scale[a] = ext[a] > 0.0f ? float(1u << L.nbits[a]) / ext[a] : 0.0f;
}

// SAH collapse data of a node (Ct = Ci = 1): it becomes a leaf when it holds at most
// max_leaf triangles and SA * count <= SA + cost(left) + cost(right). Children are always
// created before their parent (leaves first, then PLOC merges in iteration order), so the
// merge that creates a node evaluates it.
struct NodeEval {
    int32_t count;
    float cost;
    int32_t minleaf;
    int32_t height;
    uint8_t collapsed;
};

__device__ __forceinline__ NodeEval LeafEval(const float lo[3], const float hi[3], int32_t k) {
    return NodeEval{1, SurfaceAreaBox(lo, hi) * 1.0f, k, 0, 1};
}

__device__ __forceinline__ NodeEval InnerEval(const float lo[3], const float hi[3], const NodeEval& l,
                                              const NodeEval& r, int max_leaf) {
    NodeEval e;
    e.count = l.count + r.count;
    const float sa = SurfaceAreaBox(lo, hi);
    const float inner = sa + l.cost + r.cost;
    e.cost = inner;
    e.collapsed = 0;
    if (e.count <= max_leaf) {
        const float leafc = sa * float(e.count);
        if (leafc <= inner) {
            e.cost = leafc;
            e.collapsed = 1;
        }
    }
    e.minleaf = l.minleaf < r.minleaf ? l.minleaf : r.minleaf;
    e.height = 1 + (l.height > r.height ? l.height : r.height);
    return e;
}

__device__ __forceinline__ NodeEval LoadEval(const BuildBuffers& b, int32_t id) {
    return NodeEval{b.count[id], b.cost[id], b.minleaf[id], b.height[id], b.collapsed[id]};
}

__device__ __forceinline__ void StoreEval(const BuildBuffers& b, int32_t id, const NodeEval& e) {
    b.count[id] = e.count;
    b.cost[id] = e.cost;
    b.minleaf[id] = e.minleaf;
    b.height[id] = e.height;
    b.collapsed[id] = e.collapsed;
}

// ---------------------------------------------------------------------------------------
// SOBB support: k-DOP projections (n . v, FP32, left to right) padded by 1e-6 x the largest
// scene coordinate, exact unions, and the direction triple of least parallelepiped area
// (w_a w_b + w_b w_c + w_a w_c) / |det| per child box (ties: first triple).
//
// Stored k-DOP per node: only the non-axis directions 3..ndirs-1 (the axis slabs are the
// node's box -/+ pad, exactly: an axis projection is the coordinate itself and rounding
// of x - pad is monotone), as [lo_3 .. lo_nd-1, -hi_3 .. -hi_nd-1] so that a union is one
// element-wise min over 2 (ndirs - 3) floats (20 for the 13-direction set, a multiple of 4).
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ float KdopPad(const uint32_t* bounds) {
    const DecodedBounds d = Decode(bounds);
    float s = 0.0f;
    for (int a = 0; a < 3; a++) s = Max2(s, Max2(fabsf(d.vmin[a]), fabsf(d.vmax[a])));
    return Max2(1e-6f * s, 1.40129846e-45f);  // never zero: x - pad is monotone and sign-exact
}

__device__ __forceinline__ int KdopWidth(int ndirs) { return 2 * (ndirs - 3); }

__device__ __forceinline__ void LeafKdop(const BuildBuffers& b, const float v[9], float pad, float* kd) {
    const int nx = b.ndirs - 3;
    for (int k = 0; k < nx; k++) {
        const float* n = b.dirs + 3 * (3 + k);
        const float nx_ = n[0], ny = n[1], nz = n[2];
        const float p0 = nx_ * v[0] + ny * v[1] + nz * v[2];
        const float p1 = nx_ * v[3] + ny * v[4] + nz * v[5];
        const float p2 = nx_ * v[6] + ny * v[7] + nz * v[8];
        kd[k] = Min2(Min2(p0, p1), p2) - pad;
        kd[nx + k] = -(Max2(Max2(p0, p1), p2) + pad);
    }
}

__device__ __forceinline__ void UnionKdop(const BuildBuffers& b, int32_t id, int32_t l, int32_t r) {
    const int w4 = KdopWidth(b.ndirs) / 4;
    const float4* kl = reinterpret_cast<const float4*>(b.kdop) + int64_t(l) * w4;
    const float4* kr = reinterpret_cast<const float4*>(b.kdop) + int64_t(r) * w4;
    float4* ko = reinterpret_cast<float4*>(b.kdop) + int64_t(id) * w4;
    for (int k = 0; k < w4; k++) {
        const float4 x = kl[k], y = kr[k];
        ko[k] = make_float4(Min2(x.x, y.x), Min2(x.y, y.y), Min2(x.z, y.z), Min2(x.w, y.w));
    }
}

// The 13 unit normals of the 26-DOP (axes, face diagonals, corner diagonals) and, at compile
// time, the non-coplanar triples among them in lexicographic order (|det| >= 1e-3, FP64),
// identical to the run-time table the host uploads (checked when it is made).
constexpr float kS2 = 0.70710677f, kS3 = 0.57735026f;
constexpr float kDirs13[13][3] = {{1, 0, 0},     {0, 1, 0},       {0, 0, 1},       {kS2, kS2, 0},    {kS2, -kS2, 0},
                                  {kS2, 0, kS2}, {kS2, 0, -kS2},  {0, kS2, kS2},   {0, kS2, -kS2},   {kS3, kS3, kS3},
                                  {kS3, kS3, -kS3}, {kS3, -kS3, kS3}, {-kS3, kS3, kS3}};
struct TripleTable13 {
    int n;
    int8_t a[kMaxTriples], b[kMaxTriples], c[kMaxTriples];
    float inv_det[kMaxTriples];  // float(1 / |det|), the division in FP64 rounded once
};
constexpr TripleTable13 MakeTripleTable13() {
    TripleTable13 t{};
    for (int a = 0; a < 13; a++)
        for (int b = a + 1; b < 13; b++)
            for (int c = b + 1; c < 13; c++) {
                const double n1[3] = {kDirs13[a][0], kDirs13[a][1], kDirs13[a][2]};
                const double n2[3] = {kDirs13[b][0], kDirs13[b][1], kDirs13[b][2]};
                const double n3[3] = {kDirs13[c][0], kDirs13[c][1], kDirs13[c][2]};
                const double cx = n2[1] * n3[2] - n2[2] * n3[1];
                const double cy = n2[2] * n3[0] - n2[0] * n3[2];
                const double cz = n2[0] * n3[1] - n2[1] * n3[0];
                const double det = n1[0] * cx + n1[1] * cy + n1[2] * cz;
                const double ad = det < 0 ? -det : det;
                if (ad < 1e-3) continue;
                t.a[t.n] = int8_t(a), t.b[t.n] = int8_t(b), t.c[t.n] = int8_t(c);
                t.inv_det[t.n] = float(1.0 / ad);
                t.n++;
            }
    return t;
}
constexpr TripleTable13 kTriples13 = MakeTripleTable13();

template <int T>
__device__ __forceinline__ void ScanTriples13(const float* w, float& best, int& bt) {
    if constexpr (T < kTriples13.n) {
        constexpr int a = kTriples13.a[T], b = kTriples13.b[T], c = kTriples13.c[T];
        constexpr float inv_det = kTriples13.inv_det[T];
        const float sa = (w[a] * w[b] + w[b] * w[c] + w[a] * w[c]) * inv_det;
        if (T == 0 || sa < best) {
            best = sa;
            bt = T;
        }
        ScanTriples13<T + 1>(w, best, bt);
    }
}

// Selects arr[idx] for a run-time idx without indexing a private array.
template <int N>
__device__ __forceinline__ float SelDir(const float (&arr)[N], int idx) {
    float r = arr[0];
#pragma unroll
    for (int k = 1; k < N; k++) r = idx == k ? arr[k] : r;
    return r;
}

// Writes the SOBB (lo_a hi_a lo_b hi_b lo_c hi_c) of node `id` to f and returns a | b<<4 | c<<8.
// ND is the size of the direction set (3, 7, 9 or 13).
template <int ND>
__device__ __forceinline__ uint32_t FitSobbT(const BuildBuffers& b, const int4* triples, int32_t id, float pad,
                                             float* f) {
    constexpr int NX = ND - 3;
    const PNode nd = b.nodes[id];
    float lo[ND], hi[ND], w[ND];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lo[a] = nd.lo[a] - pad;
        hi[a] = nd.hi[a] + pad;
    }
    if constexpr (NX > 0) {
        const float4* kd = reinterpret_cast<const float4*>(b.kdop) + int64_t(id) * (NX / 2);
        float s[2 * NX];
#pragma unroll
        for (int q = 0; q < NX / 2; q++) {
            const float4 v = kd[q];
            s[4 * q] = v.x, s[4 * q + 1] = v.y, s[4 * q + 2] = v.z, s[4 * q + 3] = v.w;
        }
#pragma unroll
        for (int k = 0; k < NX; k++) {
            lo[3 + k] = s[k];
            hi[3 + k] = -s[NX + k];
        }
    }
#pragma unroll
    for (int k = 0; k < ND; k++) w[k] = hi[k] - lo[k];
    float best = 0.0f;
    int bt = 0;
    if constexpr (ND == 13) {
        ScanTriples13<0>(w, best, bt);
    } else {
        for (int t = 0; t < b.ntriples; t++) {
            const int4 T = triples[t];
            const float wa = SelDir(w, T.x), wb = SelDir(w, T.y), wc = SelDir(w, T.z);
            const float sa = (wa * wb + wb * wc + wa * wc) * __int_as_float(T.w);
            if (t == 0 || sa < best) {
                best = sa;
                bt = t;
            }
        }
    }
    const int4 T = triples[bt];
    f[0] = SelDir(lo, T.x), f[1] = SelDir(hi, T.x);
    f[2] = SelDir(lo, T.y), f[3] = SelDir(hi, T.y);
    f[4] = SelDir(lo, T.z), f[5] = SelDir(hi, T.z);
    return uint32_t(T.x) | (uint32_t(T.y) << 4) | (uint32_t(T.z) << 8);
}

__device__ __forceinline__ uint32_t FitSobb(const BuildBuffers& b, const int4* triples, int32_t id, float pad,
                                            float* f) {
    switch (b.ndirs) {
        case 13: return FitSobbT<13>(b, triples, id, pad, f);
        case 9: return FitSobbT<9>(b, triples, id, pad, f);
        case 7: return FitSobbT<7>(b, triples, id, pad, f);
        default: return FitSobbT<3>(b, triples, id, pad, f);
    }
}

// ---------------------------------------------------------------------------------------
// PLOC nearest neighbour over an SoA window in LDS: argmin of (SA(union), PairHash, j)
// ---------------------------------------------------------------------------------------
struct WindowLds {
    float* lx;
    float* ly;
    float* lz;
    float* hx;
    float* hy;
    float* hz;
};

__device__ __forceinline__ int32_t NearestNeighbour(const WindowLds& w, int i, int base, int c, int R) {
    const int ii = i - base;
    const float alx = w.lx[ii], aly = w.ly[ii], alz = w.lz[ii], ahx = w.hx[ii], ahy = w.hy[ii], ahz = w.hz[ii];
    const int jlo = i - R < 0 ? 0 : i - R;
    const int jhi = i + R > c - 1 ? c - 1 : i + R;
    float best_d = 0.0f;
    int32_t best_j = -1;
    uint32_t best_h = 0;
    bool hv = false;
    for (int j = jlo; j <= jhi; j++) {
        if (j == i) continue;
        const int jj = j - base;
        float d = SurfaceAreaUnion(alx, aly, alz, ahx, ahy, ahz, w.lx[jj], w.ly[jj], w.lz[jj], w.hx[jj], w.hy[jj],
                                   w.hz[jj]);
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
    return best_j;
}

// ---------------------------------------------------------------------------------------
// layout: classify, scans, emit
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ bool Inside(const BuildBuffers& b, int32_t id, int max_leaf) {
    int32_t q = b.parent[id];
    while (q >= 0 && b.count[q] <= max_leaf) {
        if (b.collapsed[q]) return true;
        q = b.parent[q];
    }
    return false;
}

// SOBB area 2 (wa wb + wb wc + wa wc) / |det(n_a, n_b, n_c)| in FP64.
__device__ inline double SobbAreaF64(const float* f, uint32_t d, const float* dirs) {
    const float* n1 = dirs + 3 * (d & 15);
    const float* n2 = dirs + 3 * ((d >> 4) & 15);
    const float* n3 = dirs + 3 * ((d >> 8) & 15);
    const double det = double(n1[0]) * (double(n2[1]) * n3[2] - double(n2[2]) * n3[1]) +
                       double(n1[1]) * (double(n2[2]) * n3[0] - double(n2[0]) * n3[2]) +
                       double(n1[2]) * (double(n2[0]) * n3[1] - double(n2[1]) * n3[0]);
    const double wa = double(f[1]) - f[0], wb = double(f[3]) - f[2], wc = double(f[5]) - f[4];
    return 2.0 * (wa * wb + wb * wc + wa * wc) / fabs(det);
}

// ---------------------------------------------------------------------------------------
// codes of one triangle in a scene (or mesh) whose bounds are d
// ---------------------------------------------------------------------------------------
struct CodeFrame {
    DecodedBounds d;
    float scale30[3];  // Morton30: 1024 / centroid extent (0 for a zero extent)
    float s2;          // EMC60: squared diagonal of the vertex bounds
};

__device__ __forceinline__ CodeFrame MakeCodeFrame(const DecodedBounds& d, int morton_type) {
    CodeFrame f;
    f.d = d;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float ext = d.cmax[a] - d.cmin[a];
        f.scale30[a] = ext > 0.0f ? 1024.0f / ext : 0.0f;
    }
    f.s2 = 0.0f;
    if (morton_type == 1) {
        const float sx = d.vmax[0] - d.vmin[0], sy = d.vmax[1] - d.vmin[1], sz = d.vmax[2] - d.vmin[2];
        f.s2 = sx * sx + sy * sy + sz * sz;
    }
    return f;
}

// Morton30: 10 bits per axis, x in the top bit of each triple. EMC60: 12 groups of [one size
// bit, four spatial bits]; spatial bits follow the adaptive allocation L, size bits are the
// float bits of diag^2 / scene diag^2 above 2^-40, shifted right by 17 and clamped to 12 bits.
template <int kMorton>
__device__ __forceinline__ uint64_t CodeOf(const float v[9], const CodeFrame& f, const EmcLayoutDev& L) {
    float c[3];
    Centroid(v, c);
    if (kMorton == 0) {
        const uint32_t q0 = Quantize(c[0], f.d.cmin[0], f.scale30[0], 1023.0f);
        const uint32_t q1 = Quantize(c[1], f.d.cmin[1], f.scale30[1], 1023.0f);
        const uint32_t q2 = Quantize(c[2], f.d.cmin[2], f.scale30[2], 1023.0f);
        return (ExpandBits10(q0) << 2) | (ExpandBits10(q1) << 1) | ExpandBits10(q2);
    }
    uint32_t q[3];
    int rem[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float maxq = float((1u << L.nbits[a]) - 1u);
        q[a] = L.nbits[a] > 0 ? Quantize(c[a], f.d.cmin[a], L.scale[a], maxq) : 0u;
        rem[a] = L.nbits[a];
    }
    float lo[3], hi[3];
    TriBox(v, lo, hi);
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float r = f.s2 > 0.0f ? d2 / f.s2 : 0.0f;
    const uint32_t bits = __float_as_uint(r);
    uint32_t sz = 0;
    if (bits > 0x2B800000u) {
        sz = (bits - 0x2B800000u) >> 17;
        if (sz > 4095u) sz = 4095u;
    }
    uint64_t code = 0;
    for (int g = 0; g < 12; g++) {
        code = (code << 1) | ((sz >> (11 - g)) & 1u);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int a = L.axis_of[4 * g + k];
            uint64_t bit = 0;
            if (a >= 0) {
                const int r0 = a == 0 ? --rem[0] : (a == 1 ? --rem[1] : --rem[2]);
                const uint32_t qa = a == 0 ? q[0] : (a == 1 ? q[1] : q[2]);
                bit = (qa >> r0) & 1u;
            }
            code = (code << 1) | bit;
        }
    }
    return code;
}

// ---------------------------------------------------------------------------------------
// leaves in sorted order: node k = triangle vals[k]; cluster k = (box, node k)
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void InitLeaf(const BuildBuffers& b, int64_t k, float pad) {
    const uint32_t orig = b.vals[k];
    float v[9];
    LoadTri(b.tri, orig, v);
    PNode nd;
    TriBox(v, nd.lo, nd.hi);
    nd.left = int32_t(orig);
    nd.right = -1;
    b.nodes[k] = nd;
    StoreEval(b, int32_t(k), LeafEval(nd.lo, nd.hi, int32_t(k)));
    if (b.ndirs > 3) LeafKdop(b, v, pad, b.kdop + k * KdopWidth(b.ndirs));
    nd.left = int32_t(k);
    nd.right = 0;
    b.clusters[0][k] = nd;
}

// ---------------------------------------------------------------------------------------
// final layout. A final leaf is a collapsed node with no collapsed ancestor within the
// max_leaf range; a kept inner node is an uncollapsed node that is not inside a final leaf.
// Kept inner nodes are numbered by descending PLOC id (the root is 0); final leaves take
// consecutive triangle ranges ordered by their smallest sorted position (minleaf).
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void ClassifyNode(const BuildBuffers& b, int32_t id, int max_leaf, int& kept,
                                             int& leaves) {
    const int64_t n = b.n;
    const bool in = Inside(b, id, max_leaf);
    const bool col = b.collapsed[id] != 0;
    if (col && !in) {
        b.leafslot[b.minleaf[id]] = b.count[id];
        leaves++;
    }
    if (id >= n) {
        const bool k = !col && !in;
        b.keptflag[id - n] = k ? 1 : 0;
        kept += k;
    }
}

// keptflag holds exclusive ranks and leafslot begin offsets (both scanned); K kept nodes.
__device__ __forceinline__ int32_t ChildRef(const BuildBuffers& b, int32_t id, int max_leaf, int32_t K) {
    const int64_t n = b.n;
    if (id >= n && !b.collapsed[id] && !Inside(b, id, max_leaf)) return K - 1 - b.keptflag[id - n];
    return ~((b.leafslot[b.minleaf[id]] << 4) | b.count[id]);
}

// Gathers the primitives of a final leaf (left-first DFS of its PLOC subtree) and writes the
// node of a kept inner node; s_tri is the SOBB triple table (LDS), pad the k-DOP pad.
__device__ __forceinline__ void EmitNode(const BuildBuffers& b, int32_t id, int max_leaf, int32_t K,
                                         const int4* s_tri, float pad) {
    const int64_t n = b.n, nn = 2 * n - 1;
    const bool col = b.collapsed[id] != 0;
    const bool in = Inside(b, id, max_leaf);
    if (col && !in) {
        int32_t pos = b.leafslot[b.minleaf[id]];
        int32_t st[kMaxLeaf + 1];
        int sp = 0;
        st[sp++] = id;
        while (sp > 0) {
            const int32_t x = st[--sp];
            if (x < n) {
                const int32_t orig = b.nodes[x].left;
                b.prim_order[pos] = orig;
                const float* p = b.tri + 9 * int64_t(orig);
                float4* d = b.tri_ordered + 3 * int64_t(pos);
                d[0] = make_float4(Canon(p[0]), Canon(p[1]), Canon(p[2]), __int_as_float(orig));
                d[1] = make_float4(Canon(p[3]), Canon(p[4]), Canon(p[5]), 0.0f);
                d[2] = make_float4(Canon(p[6]), Canon(p[7]), Canon(p[8]), 0.0f);
                pos++;
            } else if (sp + 2 <= kMaxLeaf + 1) {
                st[sp++] = b.nodes[x].right;
                st[sp++] = b.nodes[x].left;
            }
        }
        if (id == nn - 1) {  // the whole scene is one leaf: one node, empty right slot
            const PNode r = b.nodes[id];
            FNode f;
            f.lx0 = f.rx0 = r.lo[0];
            f.lx1 = f.rx1 = r.hi[0];
            f.ly0 = f.ry0 = r.lo[1];
            f.ly1 = f.ry1 = r.hi[1];
            f.lz0 = f.rz0 = r.lo[2];
            f.lz1 = f.rz1 = r.hi[2];
            f.left = ~((b.leafslot[b.minleaf[id]] << 4) | b.count[id]);
            f.right = ~0;
            f.pad0 = f.pad1 = 0;
            if (b.ndirs > 0) {
                float g[6];
                const uint32_t d = FitSobb(b, s_tri, id, pad, g);
                float* fl = &f.lx0;
                for (int k = 0; k < 6; k++) fl[k] = fl[6 + k] = g[k];
                f.pad0 = int32_t(d | (d << 12));
            }
            b.fnodes[0] = f;
        }
    } else if (id >= n && !col && !in) {
        const PNode nd = b.nodes[id];
        const int32_t left = ChildRef(b, nd.left, max_leaf, K);
        const int32_t right = ChildRef(b, nd.right, max_leaf, K);
        FNode* out = b.fnodes + (K - 1 - b.keptflag[id - n]);
        if (b.ndirs > 0) {
            // one child at a time: the 26-DOP triple scan keeps its pair products in registers
            uint32_t dirs = 0;
#pragma nounroll
            for (int c = 0; c < 2; c++) {
                float g[6];
                dirs |= FitSobb(b, s_tri, c ? nd.right : nd.left, pad, g) << (12 * c);
                float2* o = reinterpret_cast<float2*>(reinterpret_cast<float*>(out) + 6 * c);
                o[0] = make_float2(g[0], g[1]), o[1] = make_float2(g[2], g[3]), o[2] = make_float2(g[4], g[5]);
            }
            reinterpret_cast<int4*>(out)[3] = make_int4(left, right, int32_t(dirs), 0);
        } else {
            const PNode L = b.nodes[nd.left], R = b.nodes[nd.right];
            FNode f;
            f.lx0 = L.lo[0], f.lx1 = L.hi[0], f.ly0 = L.lo[1], f.ly1 = L.hi[1];
            f.rx0 = R.lo[0], f.rx1 = R.hi[0], f.ry0 = R.lo[1], f.ry1 = R.hi[1];
            f.lz0 = L.lo[2], f.lz1 = L.hi[2], f.rz0 = R.lo[2], f.rz1 = R.hi[2];
            f.left = left, f.right = right;
            f.pad0 = f.pad1 = 0;
            *out = f;
        }
    }
}

// The SAH terms (Ct = Ci = 1) one final node contributes: its children's areas, a leaf
// child's times its triangle count. AABB areas, or SOBB areas when dirs is given.
__device__ __forceinline__ double SahOfNode(const FNode& f, const float* dirs) {
    auto sa = [](float x0, float x1, float y0, float y1, float z0, float z1) {
        const double sx = double(x1) - x0, sy = double(y1) - y0, sz = double(z1) - z0;
        return 2.0 * (sx * sy + sy * sz + sx * sz);
    };
    double sl, sr;
    if (dirs) {
        sl = SobbAreaF64(&f.lx0, uint32_t(f.pad0) & 0xFFFu, dirs);
        sr = SobbAreaF64(&f.lx0 + 6, (uint32_t(f.pad0) >> 12) & 0xFFFu, dirs);
    } else {
        sl = sa(f.lx0, f.lx1, f.ly0, f.ly1, f.lz0, f.lz1);
        sr = sa(f.rx0, f.rx1, f.ry0, f.ry1, f.rz0, f.rz1);
    }
    return (f.left >= 0 ? sl : sl * double((~f.left) & 15)) + (f.right >= 0 ? sr : sr * double((~f.right) & 15));
}

// ---------------------------------------------------------------------------------------
// PLOC with every cluster resident in LDS: one workgroup of THREADS runs the remaining
// iterations on c <= CAP clusters (boxes, node ids and collapse records in LDS); merges write
// nodes, parents and collapse records to global memory. Returns the iterations done; logs
// the cluster count of every iteration to ccount_log[k] when it is given. An iteration
// without a merge cannot happen for a symmetric distance: it sets bit 1 of *flags and stops.
// ---------------------------------------------------------------------------------------
struct TailLds {
    float *lx, *ly, *lz, *hx, *hy, *hz;
    int32_t *id, *nn, *cnt, *ml, *ht;
    float* cost;
    uint32_t* scan;  // THREADS / 64 + 1 words
};

template <int THREADS, int CAP>
__device__ int TailIterations(const BuildBuffers& b, const PNode* in, int c, int R, int max_leaf, const TailLds& L,
                              int32_t* ccount_log, uint32_t* flags) {
    constexpr int IPT = (CAP + THREADS - 1) / THREADS;
    const int64_t n = b.n;
    for (int k = threadIdx.x; k < c; k += THREADS) {
        const float4* p = reinterpret_cast<const float4*>(in + k);
        const float4 q0 = p[0], q1 = p[1];
        L.lx[k] = q0.x, L.ly[k] = q0.y, L.lz[k] = q0.z, L.hx[k] = q0.w;
        L.hy[k] = q1.x, L.hz[k] = q1.y, L.id[k] = __float_as_int(q1.z);
        const NodeEval e = LoadEval(b, L.id[k]);
        L.cnt[k] = e.count, L.cost[k] = e.cost, L.ml[k] = e.minleaf, L.ht[k] = e.height;
    }
    __syncthreads();
    const WindowLds w{L.lx, L.ly, L.lz, L.hx, L.hy, L.hz};
    int it = 0;
    while (c > 1) {
        if (ccount_log && threadIdx.x == 0) ccount_log[it] = c;
        for (int i = threadIdx.x; i < c; i += THREADS) L.nn[i] = NearestNeighbour(w, i, 0, c, R);
        __syncthreads();
        uint32_t sflags = 0, mflags = 0;
        int32_t nnj[IPT];
        PNode keep[IPT];
        NodeEval keep_e[IPT];
#pragma unroll
        for (int q = 0; q < IPT; q++) {
            const int i = threadIdx.x * IPT + q;
            nnj[q] = -1;
            if (i < c) {
                const int j = L.nn[i];
                const bool mutual = L.nn[j] == i;
                nnj[q] = j;
                if (!(mutual && j < i)) sflags |= 1u << q;
                if (mutual && i < j) mflags |= 1u << q;
            }
        }
        const uint32_t mine = uint32_t(__popc(sflags)) | (uint32_t(__popc(mflags)) << 16);
        uint32_t total;
        const uint32_t before = BlockExclusiveScan<THREADS>(mine, L.scan, total);
        if ((total >> 16) == 0u) {
            if (threadIdx.x == 0) atomicOr(flags, 2u);
            break;
        }
        const int32_t node_base = int32_t(2 * n - c);
        int32_t spos = int32_t(before & 0xFFFFu);
        int32_t mpos = node_base + int32_t(before >> 16);
        // read everything this thread keeps before anyone overwrites the LDS lists
#pragma unroll
        for (int q = 0; q < IPT; q++) {
            const int i = threadIdx.x * IPT + q;
            if (!(sflags & (1u << q))) continue;
            PNode& o = keep[q];
            const NodeEval ei{L.cnt[i], L.cost[i], L.ml[i], L.ht[i], 0};
            if (mflags & (1u << q)) {
                const int j = nnj[q];
                o.lo[0] = Min2(L.lx[i], L.lx[j]);
                o.lo[1] = Min2(L.ly[i], L.ly[j]);
                o.lo[2] = Min2(L.lz[i], L.lz[j]);
                o.hi[0] = Max2(L.hx[i], L.hx[j]);
                o.hi[1] = Max2(L.hy[i], L.hy[j]);
                o.hi[2] = Max2(L.hz[i], L.hz[j]);
                const int32_t id = mpos++;
                o.left = L.id[i];
                o.right = L.id[j];
                b.nodes[id] = o;
                b.parent[o.left] = id;
                b.parent[o.right] = id;
                const NodeEval ej{L.cnt[j], L.cost[j], L.ml[j], L.ht[j], 0};
                keep_e[q] = InnerEval(o.lo, o.hi, ei, ej, max_leaf);
                StoreEval(b, id, keep_e[q]);
                if (b.ndirs > 3) UnionKdop(b, id, o.left, o.right);
                o.left = id;
            } else {
                o.lo[0] = L.lx[i], o.lo[1] = L.ly[i], o.lo[2] = L.lz[i];
                o.hi[0] = L.hx[i], o.hi[1] = L.hy[i], o.hi[2] = L.hz[i];
                o.left = L.id[i];
                keep_e[q] = ei;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < IPT; q++) {
            if (!(sflags & (1u << q))) continue;
            const PNode& o = keep[q];
            const int k = spos++;
            L.lx[k] = o.lo[0], L.ly[k] = o.lo[1], L.lz[k] = o.lo[2];
            L.hx[k] = o.hi[0], L.hy[k] = o.hi[1], L.hz[k] = o.hi[2];
            L.id[k] = o.left;
            L.cnt[k] = keep_e[q].count, L.cost[k] = keep_e[q].cost, L.ml[k] = keep_e[q].minleaf;
            L.ht[k] = keep_e[q].height;
        }
        c = int(total & 0xFFFFu);
        it++;
        __syncthreads();
    }
    if (ccount_log && threadIdx.x == 0) ccount_log[it] = c;
    return it;
}

}  // namespace dpl
