// Device helpers shared by the single-scene build kernels (build.hip) and the batched
// per-mesh build (batch.hip): triangle loads, bounds decoding, Morton30 / EMC60 codes, the
// fused leaf-collapse record, SOBB k-DOPs and fits, the PLOC nearest-neighbour search and the
// final-layout node classification. Every function follows the oracle's restatement
// operation by operation (no FMA contraction, IEEE denormals, Min2/Max2 select semantics).
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

#include "dpl_internal.h"

namespace dpl {

__device__ __forceinline__ void LoadTri(const float* tri, int64_t i, float v[9]) {
    const float* p = tri + 9 * i;
#pragma unroll
    for (int k = 0; k < 9; k++) v[k] = Canon(p[k]);
}

// LoadTri for a gather at a random index: three 16-byte loads of the aligned 48 bytes that
// hold the triangle instead of nine 4-byte loads, a third of the cache requests when every
// lane reads a different line. Only for i <= n - 2 with tri 16-byte aligned (the window then
// ends at most at float 9 n - 6, inside the array); the last triangle takes LoadTri.
__device__ __forceinline__ void GatherTri(const float* tri, int64_t i, int64_t n, bool aligned, float v[9]) {
    if (!aligned || i >= n - 1) {
        LoadTri(tri, i, v);
        return;
    }
    const int64_t base = 9 * i;
    const int64_t a = base & ~int64_t(3);
    const float4* p = reinterpret_cast<const float4*>(tri + a);
    const float4 q0 = p[0], q1 = p[1], q2 = p[2];
    const float w[12] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
    const int o = int(base - a);  // 0..3
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const float x = o == 0 ? w[k] : (o == 1 ? w[k + 1] : (o == 2 ? w[k + 2] : w[k + 3]));
        v[k] = Canon(x);
    }
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
    // deposit tables: the code bits that nibble j (from the top) of axis a's left-aligned
    // 24-bit quantised coordinate sets, for each of its 16 values; the same for the three
    // nibbles of the 12 size bits. A code is the OR of 21 table entries.
    uint64_t dep[3][6][16];
    uint64_t dep_size[3][16];
    int nbits[3];
    float scale[3];
    int8_t axis_of[48];
    int8_t pos[3][24];  // code bit (0 = least significant) of axis a's j-th bit from the top
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
            // spatial bit t sits in group t / 4 (five code bits from bit 59 down: the size bit,
            // then four spatial bits) at place t % 4
            L.pos[best][L.nbits[best]] = int8_t(58 - 5 * (t / 4) - (t % 4));
            L.nbits[best]++;
            cell[best] = cell[best] * 0.5f;
        }
    }
    for (int a = 0; a < 3; a++) L.scale[a] = ext[a] > 0.0f ? float(1u << L.nbits[a]) / ext[a] : 0.0f;
}

// The deposit tables of L (after MakeEmcLayout), built by threads tid = 0 .. nthreads - 1.
__device__ inline void BuildEmcTables(EmcLayoutDev& L, int tid, int nthreads) {
    for (int e = tid; e < 3 * 6 * 16 + 3 * 16; e += nthreads) {
        uint64_t v = 0;
        if (e < 3 * 6 * 16) {
            const int a = e / 96, j = (e / 16) % 6, x = e % 16;
            for (int i = 0; i < 4; i++) {
                const int bit = 4 * j + i;  // the axis's bit, counted from its top bit
                if (bit < L.nbits[a] && ((x >> (3 - i)) & 1)) v |= 1ull << L.pos[a][bit];
            }
            L.dep[a][j][x] = v;
        } else {
            const int k = e - 3 * 6 * 16, s = k / 16, x = k % 16;
            for (int i = 0; i < 4; i++) {
                const int g = 4 * s + i;  // size bit 11 - g goes to code bit 59 - 5 g
                if ((x >> (3 - i)) & 1) v |= 1ull << (59 - 5 * g);
            }
            L.dep_size[s][x] = v;
        }
    }
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
    const EvalRec r = b.eval[id];
    return NodeEval{r.count_col & 0x7FFFFFFF, r.cost, r.minleaf, r.height, uint8_t(uint32_t(r.count_col) >> 31)};
}

// The collapse record of node id whose box is (lo, hi). Single-scene builds store no leaf
// records: a leaf's (id < n) follows from its box and position, as LeafEval gives it.
__device__ __forceinline__ NodeEval EvalOf(const BuildBuffers& b, int32_t id, const float lo[3], const float hi[3]) {
    return id < b.n ? LeafEval(lo, hi, id) : LoadEval(b, id);
}

__device__ __forceinline__ void StoreEval(const BuildBuffers& b, int32_t id, const NodeEval& e) {
    b.eval[id] = EvalRec{int32_t(uint32_t(e.count) | (uint32_t(e.collapsed) << 31)), e.cost, e.minleaf, e.height};
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

template <int NX>
__device__ __forceinline__ void LeafKdopT(const BuildBuffers& b, const float v[9], float pad, float* kd) {
    float o[2 * NX];
#pragma unroll
    for (int k = 0; k < NX; k++) {
        const float* n = b.dirs + 3 * (3 + k);
        const float nx_ = n[0], ny = n[1], nz = n[2];
        const float p0 = nx_ * v[0] + ny * v[1] + nz * v[2];
        const float p1 = nx_ * v[3] + ny * v[4] + nz * v[5];
        const float p2 = nx_ * v[6] + ny * v[7] + nz * v[8];
        o[k] = Min2(Min2(p0, p1), p2) - pad;
        o[NX + k] = -(Max2(Max2(p0, p1), p2) + pad);
    }
    float4* d = reinterpret_cast<float4*>(kd);  // 2 NX is a multiple of 4: whole 16-byte stores
#pragma unroll
    for (int q = 0; q < NX / 2; q++) d[q] = make_float4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
}

__device__ __forceinline__ void LeafKdop(const BuildBuffers& b, const float v[9], float pad, float* kd) {
    switch (b.ndirs - 3) {
        case 10: LeafKdopT<10>(b, v, pad, kd); break;
        case 6: LeafKdopT<6>(b, v, pad, kd); break;
        case 4: LeafKdopT<4>(b, v, pad, kd); break;
        default: break;
    }
}

// All 2 W4 loads are issued before the first store: the output may alias nothing the loads
// read, but the compiler cannot know that, and interleaving would make every float4 a
// dependent round trip to memory.
template <int W4>
__device__ __forceinline__ void UnionKdopT(const BuildBuffers& b, int32_t id, int32_t l, int32_t r) {
    const float4* kl = reinterpret_cast<const float4*>(b.kdop) + int64_t(l) * W4;
    const float4* kr = reinterpret_cast<const float4*>(b.kdop) + int64_t(r) * W4;
    float4 x[W4], y[W4];
#pragma unroll
    for (int k = 0; k < W4; k++) x[k] = kl[k];
#pragma unroll
    for (int k = 0; k < W4; k++) y[k] = kr[k];
    float4* ko = reinterpret_cast<float4*>(b.kdop) + int64_t(id) * W4;
#pragma unroll
    for (int k = 0; k < W4; k++)
        ko[k] = make_float4(Min2(x[k].x, y[k].x), Min2(x[k].y, y[k].y), Min2(x[k].z, y[k].z), Min2(x[k].w, y[k].w));
}

__device__ __forceinline__ void UnionKdop(const BuildBuffers& b, int32_t id, int32_t l, int32_t r) {
    switch (KdopWidth(b.ndirs) / 4) {
        case 5: UnionKdopT<5>(b, id, l, r); break;  // 13 directions
        case 3: UnionKdopT<3>(b, id, l, r); break;  // 9
        case 2: UnionKdopT<2>(b, id, l, r); break;  // 7
        default: break;
    }
}

// The 13 unit normals of the 26-DOP (axes, face diagonals, corner diagonals) and, at compile
// time, the non-coplanar triples among them in lexicographic order (|det| >= 1e-3, FP64),
// identical to the run-time table the host uploads (checked when it is made).
constexpr float kS2 = 0.70710677f, kS3 = 0.57735026f;
constexpr float kDirs13[13][3] = {{1, 0, 0},     {0, 1, 0},       {0, 0, 1},       {kS2, kS2, 0},    {kS2, -kS2, 0},
                                  {kS2, 0, kS2}, {kS2, 0, -kS2},  {0, kS2, kS2},   {0, kS2, -kS2},   {kS3, kS3, kS3},
                                  {kS3, kS3, -kS3}, {kS3, -kS3, kS3}, {-kS3, kS3, kS3}};
struct alignas(16) Triple13 {
    int32_t a, b, c;
    float inv_det;  // float(1 / |det|), the division in FP64 rounded once
};
struct TripleTable13 {
    int n;
    Triple13 t[kMaxTriples];
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
                t.t[t.n] = Triple13{a, b, c, float(1.0 / ad)};
                t.n++;
            }
    return t;
}
constexpr TripleTable13 kTriples13 = MakeTripleTable13();

// The 26-DOP triple scan unrolled at compile time. A triple's area estimate is
// ((p_ab + p_bc) + p_ac) / |det| with p_xy = w_x w_y (x < y), the 78 pairwise width products
// computed once and shared by the triples. The triples are taken in chunks of kChunk13 in
// table order: each chunk's minimum area (fminf: the areas are finite or +inf, never NaN,
// since every width is at least twice the k-DOP pad), and the first chunk whose minimum is
// strictly smaller than every earlier chunk's. The first minimum of the whole table is the
// first triple of that chunk whose area equals its minimum (ResolveChunk13), the same triple
// as a strict-minimum scan over all triples.
constexpr int kChunk13 = 8;
constexpr int kChunks13 = (kTriples13.n + kChunk13 - 1) / kChunk13;

template <int T>
__device__ __forceinline__ float TripleArea13(const float (&p)[13][13]) {
    constexpr Triple13 tr = kTriples13.t[T];
    return ((p[tr.a][tr.b] + p[tr.b][tr.c]) + p[tr.a][tr.c]) * tr.inv_det;
}

template <int C, int U>
__device__ __forceinline__ void ChunkStep13(const float (&p)[13][13], float& m) {
    constexpr int T = C * kChunk13 + U;
    if constexpr (U > 0 && T < kTriples13.n) m = fminf(m, TripleArea13<T>(p));
}

template <int C, int... U>
__device__ __forceinline__ void ScanChunk13(const float (&p)[13][13], float& best, int& bc,
                                            std::integer_sequence<int, U...>) {
    float m = TripleArea13<C * kChunk13>(p);
    (ChunkStep13<C, U>(p, m), ...);
    if (C == 0 || m < best) {
        best = m;
        bc = C;
    }
}

template <int... C>
__device__ __forceinline__ void ScanChunks13(const float (&p)[13][13], float& best, int& bc,
                                             std::integer_sequence<int, C...>) {
    (ScanChunk13<C>(p, best, bc, std::make_integer_sequence<int, kChunk13>{}), ...);
}

// The first triple of chunk bc whose area equals best (the chunk's minimum), from the slabs in
// this lane's LDS column wl (lo_k at k, hi_k at 13 + k, stride apart; width = hi - lo, the
// same subtraction as the scan's) and the triple table in LDS (the same triples in the same
// order, with the same 1/|det| bits): the same operations give the same areas.
__device__ __forceinline__ uint32_t ResolveChunk13(const float* wl, int stride, const int4* triples, int bc,
                                                   float best) {
    uint32_t abc = 0;
#pragma unroll
    for (int u = kChunk13 - 1; u >= 0; u--) {
        const int t = bc * kChunk13 + u;
        if (t < kTriples13.n) {
            const int4 T = triples[t];
            const float wa = wl[(13 + T.x) * stride] - wl[T.x * stride];
            const float wb = wl[(13 + T.y) * stride] - wl[T.y * stride];
            const float wc = wl[(13 + T.z) * stride] - wl[T.z * stride];
            const float sa = ((wa * wb + wb * wc) + wa * wc) * __int_as_float(T.w);
            if (sa == best) abc = uint32_t(T.x) | (uint32_t(T.y) << 4) | (uint32_t(T.z) << 8);
        }
    }
    return abc;
}

// Writes the SOBB (lo_a hi_a lo_b hi_b lo_c hi_c) of node `id` to f and returns a | b<<4 | c<<8.
// ND is the size of the direction set (3, 7, 9 or 13).
// Slab s of direction k of node id: axis directions from the node box -/+ pad, the others
// from the stored k-DOP slabs (lo, -hi).
__device__ __forceinline__ void SlabOf(const BuildBuffers& b, const PNode& nd, int32_t id, int k, float pad, float& lo,
                                       float& hi) {
    if (k < 3) {
        // selects between named values: an index into the node's arrays would put them in
        // scratch memory
        const float l0 = nd.lo[0], l1 = nd.lo[1], l2 = nd.lo[2], h0 = nd.hi[0], h1 = nd.hi[1], h2 = nd.hi[2];
        const float l = k == 0 ? l0 : (k == 1 ? l1 : l2);
        const float h = k == 0 ? h0 : (k == 1 ? h1 : h2);
        lo = l - pad;
        hi = h + pad;
    } else {
        const int nx = b.ndirs - 3;
        const float* kd = b.kdop + int64_t(id) * (2 * nx);
        lo = kd[k - 3];
        hi = -kd[nx + k - 3];
    }
}

// Writes the SOBB (lo_a hi_a lo_b hi_b lo_c hi_c) of node `id` to f and returns a | b<<4 | c<<8.
// ND is the size of the direction set (3, 7, 9 or 13). The widths go to this thread's LDS
// column wl (stride apart) and the triple scan reads them there: register arrays indexed at
// run time would spill.
template <int ND>
__device__ __forceinline__ uint32_t FitSobbT(const BuildBuffers& b, const int4* triples, int32_t id, float pad,
                                             float* f, float* wl, int stride) {
    constexpr int NX = ND - 3;
    const PNode nd = b.nodes[id];
    if constexpr (ND == 13) {
        // widths and slabs in registers: [lo_3..lo_12, -hi_3..-hi_12] after the three axes
        const float4* kd = reinterpret_cast<const float4*>(b.kdop) + int64_t(id) * (NX / 2);
        float s[2 * NX];
#pragma unroll
        for (int q = 0; q < NX / 2; q++) {
            const float4 v = kd[q];
            s[4 * q] = v.x, s[4 * q + 1] = v.y, s[4 * q + 2] = v.z, s[4 * q + 3] = v.w;
        }
        float lo[13], hi[13], w[13];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo[a] = nd.lo[a] - pad;
            hi[a] = nd.hi[a] + pad;
            w[a] = hi[a] - lo[a];
        }
#pragma unroll
        for (int k = 0; k < NX; k++) {
            lo[3 + k] = s[k];
            hi[3 + k] = -s[NX + k];
            w[3 + k] = -s[NX + k] - s[k];
        }
        // the slabs go to this lane's LDS column (lo at 0..12, hi at 13..25, stride apart): the
        // chunk resolution and the chosen slabs read them there
#pragma unroll
        for (int k = 0; k < 13; k++) {
            wl[k * stride] = lo[k];
            wl[(13 + k) * stride] = hi[k];
        }
        float p[13][13];
#pragma unroll
        for (int x = 0; x < 13; x++)
#pragma unroll
            for (int y = x + 1; y < 13; y++) p[x][y] = w[x] * w[y];
        float best = 0.0f;
        int bc = 0;
        ScanChunks13(p, best, bc, std::make_integer_sequence<int, kChunks13>{});
        const uint32_t abc = ResolveChunk13(wl, stride, triples, bc, best);
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const uint32_t k = (abc >> (4 * q)) & 15u;
            f[2 * q] = wl[k * stride];
            f[2 * q + 1] = wl[(13 + k) * stride];
        }
        return abc;
    }
#pragma unroll
    for (int a = 0; a < 3; a++) wl[a * stride] = (nd.hi[a] + pad) - (nd.lo[a] - pad);
    if constexpr (NX > 0) {
        const float4* kd = reinterpret_cast<const float4*>(b.kdop) + int64_t(id) * (NX / 2);
        float s[2 * NX];
#pragma unroll
        for (int q = 0; q < NX / 2; q++) {
            const float4 v = kd[q];
            s[4 * q] = v.x, s[4 * q + 1] = v.y, s[4 * q + 2] = v.z, s[4 * q + 3] = v.w;
        }
#pragma unroll
        for (int k = 0; k < NX; k++) wl[(3 + k) * stride] = -s[NX + k] - s[k];
    }
    float best = 0.0f;
    int bt = 0;
    for (int t = 0; t < b.ntriples; t++) {
        const int4 T = triples[t];
        const float wa = wl[T.x * stride], wb = wl[T.y * stride], wc = wl[T.z * stride];
        const float sa = (wa * wb + wb * wc + wa * wc) * __int_as_float(T.w);
        if (t == 0 || sa < best) {
            best = sa;
            bt = t;
        }
    }
    const int4 T = triples[bt];
    SlabOf(b, nd, id, T.x, pad, f[0], f[1]);
    SlabOf(b, nd, id, T.y, pad, f[2], f[3]);
    SlabOf(b, nd, id, T.z, pad, f[4], f[5]);
    return uint32_t(T.x) | (uint32_t(T.y) << 4) | (uint32_t(T.z) << 8);
}

// wl / stride: 2 kMaxDirs LDS floats of this thread (widths; the 26-DOP scan keeps its slabs there).
__device__ __forceinline__ uint32_t FitSobb(const BuildBuffers& b, const int4* triples, int32_t id, float pad,
                                            float* f, float* wl, int stride) {
    switch (b.ndirs) {
        case 13: return FitSobbT<13>(b, triples, id, pad, f, wl, stride);
        case 9: return FitSobbT<9>(b, triples, id, pad, f, wl, stride);
        case 7: return FitSobbT<7>(b, triples, id, pad, f, wl, stride);
        default: return FitSobbT<3>(b, triples, id, pad, f, wl, stride);
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

// key(j) < key(k) for the nearest-neighbour order (area, PairHash(i, .), index); the hashes
// are computed only on an exact area tie.
__device__ __forceinline__ bool NnBefore(int i, float dj, int32_t j, float dk, int32_t k) {
    if (dj != dk) return dj < dk;
    const uint32_t hj = PairHash(i, j), hk = PairHash(i, k);
    return hj < hk || (hj == hk && j < k);
}

// SurfaceAreaUnion for the nearest-neighbour search, with the IEEE minimum / maximum
// instructions: on finite boxes they pick the same values as the select forms except for the
// sign of a zero, which changes at most the sign of a zero area, and the search only
// compares areas (where -0 == +0).
__device__ __forceinline__ float NnArea(float alx, float aly, float alz, float ahx, float ahy, float ahz, const WindowLds& w,
                                        int jj) {
    constexpr float kMax = 3.40282347e38f;
    const float sx = __builtin_elementwise_minimum(
        __builtin_elementwise_maximum(ahx, w.hx[jj]) - __builtin_elementwise_minimum(alx, w.lx[jj]), kMax);
    const float sy = __builtin_elementwise_minimum(
        __builtin_elementwise_maximum(ahy, w.hy[jj]) - __builtin_elementwise_minimum(aly, w.ly[jj]), kMax);
    const float sz = __builtin_elementwise_minimum(
        __builtin_elementwise_maximum(ahz, w.hz[jj]) - __builtin_elementwise_minimum(alz, w.lz[jj]), kMax);
    return 2.0f * (sx * sy + sy * sz + sx * sz);
}

// One candidate of a chain: the chain's best (bd, bj) moves to (d, j) when key(j) is smaller
// (within a chain j grows, so an equal area is taken only for a smaller hash).
__device__ __forceinline__ void NnOffer(int i, int j, float d, bool valid, float& bd, int32_t& bj) {
    bool take = valid && d < bd;
    if (valid && d == bd) take = bj < 0 || PairHash(i, j) < PairHash(i, bj);
    bd = take ? d : bd;
    bj = take ? j : bj;
}

// The part of the nearest-neighbour search of cluster i over candidates j = jlo + part,
// jlo + part + step, ... (0 < |j - i| <= R): the minimum of key (SA(box_i U box_j),
// PairHash(i, j), j) in (bd, bj), bj = -1 when there is none. Four independent chains keep
// four running minima merged at the end: the key is a strict total order, so the minimum of
// the chain minima is the minimum, and the chains' LDS loads and arithmetic overlap.
__device__ __forceinline__ void NnScan(const WindowLds& w, int i, int base, int c, int R, int part, int step, float& bd,
                                       int32_t& bj) {
    const int ii = i - base;
    const float alx = w.lx[ii], aly = w.ly[ii], alz = w.lz[ii], ahx = w.hx[ii], ahy = w.hy[ii], ahz = w.hz[ii];
    const int jlo = i - R < 0 ? 0 : i - R;
    const int jhi = i + R > c - 1 ? c - 1 : i + R;
    constexpr float kInf = __builtin_huge_valf();
    float bd0 = kInf, bd1 = kInf, bd2 = kInf, bd3 = kInf;
    int32_t bj0 = -1, bj1 = -1, bj2 = -1, bj3 = -1;
    for (int j0 = jlo + part; j0 <= jhi; j0 += 4 * step) {
        const int j1 = j0 + step, j2 = j0 + 2 * step, j3 = j0 + 3 * step;
        // past the end: re-read the last candidate (in the window) and discard it
        const float d0 = NnArea(alx, aly, alz, ahx, ahy, ahz, w, j0 - base);
        const float d1 = NnArea(alx, aly, alz, ahx, ahy, ahz, w, (j1 <= jhi ? j1 : jhi) - base);
        const float d2 = NnArea(alx, aly, alz, ahx, ahy, ahz, w, (j2 <= jhi ? j2 : jhi) - base);
        const float d3 = NnArea(alx, aly, alz, ahx, ahy, ahz, w, (j3 <= jhi ? j3 : jhi) - base);
        NnOffer(i, j0, d0, j0 != i, bd0, bj0);
        NnOffer(i, j1, d1, j1 <= jhi && j1 != i, bd1, bj1);
        NnOffer(i, j2, d2, j2 <= jhi && j2 != i, bd2, bj2);
        NnOffer(i, j3, d3, j3 <= jhi && j3 != i, bd3, bj3);
    }
    bd = bd0;
    bj = bj0;
    if (bj1 >= 0 && (bj < 0 || NnBefore(i, bd1, bj1, bd, bj))) bd = bd1, bj = bj1;
    if (bj2 >= 0 && (bj < 0 || NnBefore(i, bd2, bj2, bd, bj))) bd = bd2, bj = bj2;
    if (bj3 >= 0 && (bj < 0 || NnBefore(i, bd3, bj3, bd, bj))) bd = bd3, bj = bj3;
}

// argmin over 0 < |j - i| <= R of (SA(box_i U box_j), PairHash(i, j), j).
__device__ __forceinline__ int32_t NearestNeighbour(const WindowLds& w, int i, int base, int c, int R) {
    float bd;
    int32_t bj;
    NnScan(w, i, base, c, R, 0, 1, bd, bj);
    return bj;
}

// The same search split over `lanes` adjacent lanes of a wave (1, 2 or 4; lane part =
// lane % lanes scans every lanes-th candidate), combined by shuffles: every lane of the
// group returns the neighbour. Lanes with active = false take part in the shuffles only.
__device__ __forceinline__ int32_t NearestNeighbourLanes(const WindowLds& w, int i, int base, int c, int R, int lanes,
                                                         bool active) {
    const int part = int(threadIdx.x) % lanes;
    float bd = 0.0f;
    int32_t bj = -1;
    if (active) NnScan(w, i, base, c, R, part, lanes, bd, bj);
    for (int m = 1; m < lanes; m <<= 1) {
        const float od = __shfl_xor(bd, m, kWave);
        const int32_t oj = __shfl_xor(bj, m, kWave);
        if (oj >= 0 && (bj < 0 || NnBefore(i, od, oj, bd, bj))) bd = od, bj = oj;
    }
    return bj;
}

// ---------------------------------------------------------------------------------------
// layout: classify, scans, emit
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ bool Inside(const BuildBuffers& b, int32_t id, int max_leaf) {
    int32_t q = b.parent[id];
    while (q >= 0) {
        const int32_t cc = b.eval[q].count_col;
        if ((cc & 0x7FFFFFFF) > max_leaf) break;
        if (cc < 0) return true;  // collapsed
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
    // 12 groups of [size bit, four spatial bits]: OR of the deposit-table entries of every
    // nibble (the tables place each bit where MakeEmcLayout's allocation puts it)
    uint64_t code = L.dep_size[0][(sz >> 8) & 15u] | L.dep_size[1][(sz >> 4) & 15u] | L.dep_size[2][sz & 15u];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const uint32_t top = q[a] << (24 - rem[a]);  // left-aligned in 24 bits (q = 0 when rem = 0)
#pragma unroll
        for (int j = 0; j < 6; j++) code |= L.dep[a][j][(top >> (20 - 4 * j)) & 15u];
    }
    return code;
}

// ---------------------------------------------------------------------------------------
// leaves in sorted order: node k = triangle vals[k]; cluster k = (box, node k)
// ---------------------------------------------------------------------------------------
// GATHER16: read the triangle with GatherTri (single-scene builds, where the gathers miss the
// caches); batched meshes are small and cache-resident, where the nine plain loads are cheaper.
// SINGLE (single-scene builds): only the leaf node and its k-DOP are written; the first
// iteration reads the leaves from the node array, and a leaf's collapse record is EvalOf's.
template <bool GATHER16 = false, bool SINGLE = false>
__device__ __forceinline__ void InitLeaf(const BuildBuffers& b, int64_t k, float pad) {
    const uint32_t orig = b.vals[k];
    float v[9];
    if (GATHER16)
        GatherTri(b.tri, orig, b.n, (reinterpret_cast<uintptr_t>(b.tri) & 15) == 0, v);
    else
        LoadTri(b.tri, orig, v);
    PNode nd;
    TriBox(v, nd.lo, nd.hi);
    nd.left = int32_t(orig);
    nd.right = -1;
    b.nodes[k] = nd;
    if (b.ndirs > 3) LeafKdop(b, v, pad, b.kdop + k * KdopWidth(b.ndirs));
    if (SINGLE) return;
    StoreEval(b, int32_t(k), LeafEval(nd.lo, nd.hi, int32_t(k)));
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
    // a leaf's record: one triangle, collapsed, its own position as minleaf
    const EvalRec e = id < n ? EvalRec{int32_t(1u | (1u << 31)), 0.0f, id, 0} : b.eval[id];
    // ancestors hold more triangles: from max_leaf up, none can be a collapsed one
    const bool in = (e.count_col & 0x7FFFFFFF) < max_leaf && Inside(b, id, max_leaf);
    const bool col = e.count_col < 0;
    uint8_t st = 0;
    if (col && !in) {
        b.leafslot[e.minleaf] = e.count_col & 0x7FFFFFFF;
        leaves++;
        st = 1;
    }
    if (id >= n) {
        const bool k = !col && !in;
        b.keptflag[id - n] = k ? 1 : 0;
        kept += k;
        if (k) st = 2;
    }
    b.state[id] = st;
}

// keptflag holds exclusive ranks and leafslot begin offsets (both scanned); K kept nodes.
// A child of a kept node is itself kept (state 2) or a final leaf (state 1).
__device__ __forceinline__ int32_t ChildRef(const BuildBuffers& b, int32_t id, int32_t K) {
    if (b.state[id] == 2) return K - 1 - b.keptflag[id - b.n];
    if (id < b.n) return ~((b.leafslot[id] << 4) | 1);  // one triangle: minleaf is its own position
    const EvalRec e = b.eval[id];
    return ~((b.leafslot[e.minleaf] << 4) | (e.count_col & 0x7FFFFFFF));
}

// The final position of sorted triangle k: its final leaf's first slot plus the triangles
// before it in the leaf's left-first order (the left siblings' counts on the way up).
template <bool GATHER16 = false>
__device__ __forceinline__ void EmitTriangle(const BuildBuffers& b, int32_t k) {
    int32_t x = k, off = 0;
    int32_t first = k;  // a one-triangle leaf starts at its own position's slot
    if (b.state[k] == 0) {
        do {
            const int32_t p = b.parent[x];
            const int32_t left = b.nodes[p].left;
            if (x != left) off += left < b.n ? 1 : (b.eval[left].count_col & 0x7FFFFFFF);
            x = p;
        } while (b.state[x] == 0);
        first = b.eval[x].minleaf;
    }
    const int32_t pos = b.leafslot[first] + off;
    const int32_t orig = int32_t(b.vals[k]);  // leaf k's triangle (the sorted ids: 4 bytes, not its 32-byte node)
    b.prim_order[pos] = orig;
    float p[9];
    if (GATHER16)
        GatherTri(b.tri, orig, b.n, (reinterpret_cast<uintptr_t>(b.tri) & 15) == 0, p);
    else
        LoadTri(b.tri, orig, p);
    float4* d = b.tri_ordered + 3 * int64_t(pos);
    d[0] = make_float4(p[0], p[1], p[2], __int_as_float(orig));
    d[1] = make_float4(p[3], p[4], p[5], 0.0f);
    d[2] = make_float4(p[6], p[7], p[8], 0.0f);
}

// Writes the node of a kept inner node (for SOBB trees its child references only; FitNode
// adds the child slabs), or, when the whole scene is one leaf, that leaf as the root.
__device__ __forceinline__ void EmitNode(const BuildBuffers& b, int32_t id, int32_t K) {
    const int64_t n = b.n, nn = 2 * n - 1;
    const uint8_t st = b.state[id];
    if (st == 1) {
        if (id == nn - 1) {  // the whole scene is one leaf: one node, empty right slot
            const EvalRec e = id < n ? EvalRec{int32_t(1u | (1u << 31)), 0.0f, id, 0} : b.eval[id];
            const PNode r = b.nodes[id];
            FNode f;
            f.lx0 = f.rx0 = r.lo[0];
            f.lx1 = f.rx1 = r.hi[0];
            f.ly0 = f.ry0 = r.lo[1];
            f.ly1 = f.ry1 = r.hi[1];
            f.lz0 = f.rz0 = r.lo[2];
            f.lz1 = f.rz1 = r.hi[2];
            f.left = ~((b.leafslot[e.minleaf] << 4) | (e.count_col & 0x7FFFFFFF));
            f.right = ~0;
            f.pad0 = f.pad1 = 0;  // SOBB trees: FitNode writes the slabs and directions
            b.fnodes[0] = f;
        }
    } else if (st == 2) {
        const PNode nd = b.nodes[id];
        const int32_t left = ChildRef(b, nd.left, K);
        const int32_t right = ChildRef(b, nd.right, K);
        const int32_t slot = K - 1 - b.keptflag[id - n];
        FNode* out = b.fnodes + slot;
        if (b.ndirs > 0) {  // child slabs and directions: FitNode, one thread per node
            b.kept_child[slot] = make_int2(nd.left, nd.right);
            reinterpret_cast<int4*>(out)[3] = make_int4(left, right, 0, 0);
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

// SOBBs of both children of final inner node `slot` (or, when the whole scene is one leaf,
// slot 0 gets the root's SOBB in both halves): slabs to the node's float slots, directions
// to its pad word (EmitNode wrote the child references and a zero pad word).
__device__ __forceinline__ void FitNode(const BuildBuffers& b, int64_t slot, int32_t K, const int4* s_tri, float pad,
                                        float* wl, int stride) {
    float g[6];
    float2* o = reinterpret_cast<float2*>(b.fnodes + slot);
    if (K == 0) {
        if (slot != 0) return;
        const uint32_t d = FitSobb(b, s_tri, int32_t(2 * b.n - 2), pad, g, wl, stride);
        for (int k = 0; k < 3; k++) o[k] = o[3 + k] = make_float2(g[2 * k], g[2 * k + 1]);
        b.fnodes[0].pad0 = int32_t(d | (d << 12));
        return;
    }
    const int2 ch = b.kept_child[slot];
    const uint32_t dl = FitSobb(b, s_tri, ch.x, pad, g, wl, stride);
    o[0] = make_float2(g[0], g[1]), o[1] = make_float2(g[2], g[3]), o[2] = make_float2(g[4], g[5]);
    const uint32_t dr = FitSobb(b, s_tri, ch.y, pad, g, wl, stride);
    o[3] = make_float2(g[0], g[1]), o[4] = make_float2(g[2], g[3]), o[5] = make_float2(g[4], g[5]);
    b.fnodes[slot].pad0 = int32_t(dl | (dr << 12));
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
                              int32_t* ccount_log, uint32_t* flags, bool leaf_input = false) {
    constexpr int IPT = (CAP + THREADS - 1) / THREADS;
    const int64_t n = b.n;
    for (int k = threadIdx.x; k < c; k += THREADS) {
        const float4* p = reinterpret_cast<const float4*>(in + k);
        const float4 q0 = p[0], q1 = p[1];
        L.lx[k] = q0.x, L.ly[k] = q0.y, L.lz[k] = q0.z, L.hx[k] = q0.w;
        L.hy[k] = q1.x, L.hz[k] = q1.y, L.id[k] = leaf_input ? k : __float_as_int(q1.z);
        const float lo[3] = {q0.x, q0.y, q0.z}, hi[3] = {q0.w, q1.x, q1.y};
        const NodeEval e = EvalOf(b, L.id[k], lo, hi);
        L.cnt[k] = e.count, L.cost[k] = e.cost, L.ml[k] = e.minleaf, L.ht[k] = e.height;
    }
    __syncthreads();
    const WindowLds w{L.lx, L.ly, L.lz, L.hx, L.hy, L.hz};
    int it = 0;
    while (c > 1) {
        if (ccount_log && threadIdx.x == 0) ccount_log[it] = c;
        if (c > THREADS / 2) {
            for (int i = threadIdx.x; i < c; i += THREADS) L.nn[i] = NearestNeighbour(w, i, 0, c, R);
        } else {  // spare lanes: two or four per cluster
            const int lanes = c > THREADS / 4 ? 2 : 4;
            const int i = int(threadIdx.x) / lanes;
            const int32_t j = NearestNeighbourLanes(w, i < c ? i : 0, 0, c, R, lanes, i < c);
            if (i < c && int(threadIdx.x) % lanes == 0) L.nn[i] = j;
        }
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
