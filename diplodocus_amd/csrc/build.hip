// BVH build kernels for gfx950: scene/centroid bounds, Morton30 / EMC60 codes, PLOC
// clustering (Meister & Bittner 2017) with one fused kernel per iteration (nearest
// neighbours over an LDS window, mutual-pair merge, order-preserving compaction with a
// decoupled look-back), a single-workgroup tail that finishes the last iterations in LDS,
// a bottom-up SAH leaf collapse, and the final two-child node layout.
//
// The CPU restatement in the oracle defines the exact results; see the bit-exactness
// contract in dpl_internal.h.
#include <hip/hip_runtime.h>

#include "dpl_internal.h"

namespace dpl {

namespace {

constexpr int kTail = 2048;        // clusters finished by the single-workgroup tail
constexpr int kTailThreads = 1024;
constexpr int kIpt = 4;            // clusters per thread in an iteration tile

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

}  // namespace

// ---------------------------------------------------------------------------------------
// init: bounds sentinels, counters, first look-back buffer, root parent
// ---------------------------------------------------------------------------------------
__global__ void KBuildInit(BuildBuffers b, int tiles0) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 6) {
        b.bounds[(t / 3) * 6 + (t % 3)] = 0xFFFFFFFFu;     // vmin, cmin
        b.bounds[(t / 3) * 6 + 3 + (t % 3)] = 0u;          // vmax, cmax
    }
    if (t == 6) b.bounds[12] = 0u;
    if (t == 7) {
        b.tile_counter[0] = 0u;
        b.tile_counter[1] = 0u;
        b.ccount[0] = int32_t(b.n);
        b.iter_done[0] = 0;
        b.parent[2 * b.n - 2] = -1;
        b.sah[0] = 0.0;
    }
    if (t < 8) b.counters[t] = 0;
    for (int64_t k = t; k < tiles0; k += int64_t(gridDim.x) * blockDim.x) b.status[0][k] = 0ull;
    for (int64_t k = t; k < b.n; k += int64_t(gridDim.x) * blockDim.x) b.leafslot[k] = 0;
}

hipError_t LaunchBuildInit(const BuildBuffers& b, int block, hipStream_t s) {
    const int64_t tiles0 = PlocStatusWords(b.n, block);
    int grid = int(std::min<int64_t>(1024, (std::max<int64_t>(b.n, tiles0) + 255) / 256));
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(KBuildInit, dim3(grid), dim3(256), 0, s, b, int(tiles0));
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// bounds: vertex and centroid min/max, exact in any order; non-finite input flagged
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void KBounds(const float* __restrict__ tri, int64_t n, uint32_t* bounds) {
    float mn[6], mx[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        mn[k] = __int_as_float(0x7f800000);
        mx[k] = -__int_as_float(0x7f800000);
    }
    bool bad = false;
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        float v[9];
        LoadTri(tri, i, v);
#pragma unroll
        for (int k = 0; k < 9; k++) bad |= !isfinite(v[k]);
        float lo[3], hi[3], c[3];
        TriBox(v, lo, hi);
        Centroid(v, c);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            mn[a] = fminf(mn[a], lo[a]);
            mx[a] = fmaxf(mx[a], hi[a]);
            mn[3 + a] = fminf(mn[3 + a], c[a]);
            mx[3 + a] = fmaxf(mx[3 + a], c[a]);
        }
    }
    __shared__ float s_mn[6][4], s_mx[6][4];
    __shared__ int s_bad;
    if (threadIdx.x == 0) s_bad = 0;
    const int wid = threadIdx.x / kWave, lane = LaneId();
#pragma unroll
    for (int k = 0; k < 6; k++) {
        float a = WaveMin(mn[k]), z = WaveMax(mx[k]);
        if (lane == 0) {
            s_mn[k][wid] = a;
            s_mx[k][wid] = z;
        }
    }
    __syncthreads();
    if (bad) atomicOr(&s_bad, 1);
    __syncthreads();
    if (threadIdx.x < 6) {
        int k = threadIdx.x;
        float a = s_mn[k][0], z = s_mx[k][0];
        for (int w = 1; w < int(blockDim.x / kWave); w++) {
            a = fminf(a, s_mn[k][w]);
            z = fmaxf(z, s_mx[k][w]);
        }
        // slots: vmin 0-2, vmax 3-5, cmin 6-8, cmax 9-11
        int base = k < 3 ? 0 : 6, ax = k % 3;
        atomicMin(&bounds[base + ax], FloatToOrdered(a));
        atomicMax(&bounds[base + 3 + ax], FloatToOrdered(z));
    }
    if (threadIdx.x == 0 && s_bad) atomicOr(&bounds[12], 1u);
}

hipError_t LaunchBounds(const BuildBuffers& b, hipStream_t s) {
    int grid = int(std::min<int64_t>(2048, (b.n + 255) / 256));
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(KBounds, dim3(grid), dim3(256), 0, s, b.tri, b.n, b.bounds);
    return hipGetLastError();
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
__device__ void MakeEmcLayout(const DecodedBounds& d, EmcLayoutDev& L) {
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

template <int kMorton>
__global__ __launch_bounds__(256) void KCodes(const float* __restrict__ tri, int64_t n, const uint32_t* bounds,
                                               uint64_t* keys, uint32_t* vals) {
    __shared__ EmcLayoutDev sL;
    const DecodedBounds d = Decode(bounds);
    if (kMorton == 1) {
        if (threadIdx.x == 0) MakeEmcLayout(d, sL);
        __syncthreads();
    }
    float scale30[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float ext = d.cmax[a] - d.cmin[a];
        scale30[a] = ext > 0.0f ? 1024.0f / ext : 0.0f;
    }
    float s2 = 0.0f;
    if (kMorton == 1) {
        float sx = d.vmax[0] - d.vmin[0], sy = d.vmax[1] - d.vmin[1], sz = d.vmax[2] - d.vmin[2];
        s2 = sx * sx + sy * sy + sz * sz;
    }
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        float v[9], c[3];
        LoadTri(tri, i, v);
        Centroid(v, c);
        if (kMorton == 0) {
            uint32_t q0 = Quantize(c[0], d.cmin[0], scale30[0], 1023.0f);
            uint32_t q1 = Quantize(c[1], d.cmin[1], scale30[1], 1023.0f);
            uint32_t q2 = Quantize(c[2], d.cmin[2], scale30[2], 1023.0f);
            reinterpret_cast<uint32_t*>(keys)[i] = (ExpandBits10(q0) << 2) | (ExpandBits10(q1) << 1) | ExpandBits10(q2);
        } else {
            uint32_t q[3];
            int rem[3];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                float maxq = float((1u << sL.nbits[a]) - 1u);
                q[a] = sL.nbits[a] > 0 ? Quantize(c[a], d.cmin[a], sL.scale[a], maxq) : 0u;
                rem[a] = sL.nbits[a];
            }
            float lo[3], hi[3];
            TriBox(v, lo, hi);
            float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
            float d2 = dx * dx + dy * dy + dz * dz;
            float r = s2 > 0.0f ? d2 / s2 : 0.0f;
            uint32_t bits = __float_as_uint(r);
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
                    int a = sL.axis_of[4 * g + k];
                    uint64_t bit = 0;
                    if (a >= 0) {
                        int r0 = a == 0 ? --rem[0] : (a == 1 ? --rem[1] : --rem[2]);
                        uint32_t qa = a == 0 ? q[0] : (a == 1 ? q[1] : q[2]);
                        bit = (qa >> r0) & 1u;
                    }
                    code = (code << 1) | bit;
                }
            }
            keys[i] = code;
        }
        vals[i] = uint32_t(i);
    }
}

hipError_t LaunchCodes(const BuildBuffers& b, int morton_type, hipStream_t s) {
    int grid = int(std::min<int64_t>(4096, (b.n + 255) / 256));
    if (grid < 1) grid = 1;
    if (morton_type == 0)
        hipLaunchKernelGGL(KCodes<0>, dim3(grid), dim3(256), 0, s, b.tri, b.n, b.bounds, b.keys, b.vals);
    else
        hipLaunchKernelGGL(KCodes<1>, dim3(grid), dim3(256), 0, s, b.tri, b.n, b.bounds, b.keys, b.vals);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// leaves in sorted order: node k = triangle vals[k]; cluster k = (box, node k)
// ---------------------------------------------------------------------------------------
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
int SobbTriples13() { return kTriples13.n; }
void SobbTriple13(int t, int* a, int* b, int* c, float* inv_det) {
    *a = kTriples13.a[t], *b = kTriples13.b[t], *c = kTriples13.c[t], *inv_det = kTriples13.inv_det[t];
}

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

__global__ __launch_bounds__(256) void KInitClusters(BuildBuffers b) {
    const int64_t n = b.n;
    const float pad = b.ndirs > 0 ? KdopPad(b.bounds) : 0.0f;
    for (int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < n; k += int64_t(gridDim.x) * blockDim.x) {
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
}

hipError_t LaunchInitClusters(const BuildBuffers& b, hipStream_t s) {
    int grid = int(std::min<int64_t>(4096, (b.n + 255) / 256));
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(KInitClusters, dim3(grid), dim3(256), 0, s, b);
    return hipGetLastError();
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
// one PLOC iteration over tiles of CHUNK = BLOCK * kIpt clusters; persistent over tiles
// with dynamic tile ids so the decoupled look-back always waits on running workgroups
// ---------------------------------------------------------------------------------------
constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagInc = 2ull << 62;
constexpr uint64_t kMask31 = 0x7FFFFFFFull;

// Tiles hold BLOCK * ipt clusters, ipt in {1, 2, 4} chosen per iteration so that there are at
// least `min_tiles` tiles: small iterations spread over more workgroups, each thread then
// computing about one nearest neighbour instead of four.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void KPlocIteration(BuildBuffers b, int iter, int R, int tail, int max_leaf,
                                                        int min_tiles) {
    constexpr int CHUNK = BLOCK * kIpt;
    constexpr int WMAX = CHUNK + 4 * kMaxRadius;
    __shared__ float s_lx[WMAX], s_ly[WMAX], s_lz[WMAX], s_hx[WMAX], s_hy[WMAX], s_hz[WMAX];
    __shared__ int32_t s_id[WMAX];
    __shared__ int32_t s_nn[CHUNK + 2 * kMaxRadius];
    __shared__ uint32_t s_scan[BLOCK / kWave + 1];
    __shared__ int s_tile;
    __shared__ uint64_t s_prefix;

    const int c = b.ccount[iter];
    const PNode* __restrict__ in = b.clusters[iter & 1];
    PNode* __restrict__ out = b.clusters[(iter + 1) & 1];
    const int64_t n = b.n;

    if (c <= tail) {  // done with multi-workgroup iterations: pass the clusters through
        if (blockIdx.x == 0) {
            for (int k = threadIdx.x; k < c; k += BLOCK) out[k] = in[k];
            if (threadIdx.x == 0) {
                b.ccount[iter + 1] = c;
                b.tile_counter[(iter + 1) & 1] = 0u;
                b.results->progress = c;
            }
        }
        return;
    }
    const int ipt = c >= int64_t(BLOCK) * 4 * min_tiles ? 4 : (c >= int64_t(BLOCK) * 2 * min_tiles ? 2 : 1);
    const int chunk = BLOCK * ipt;
    const int T = (c + chunk - 1) / chunk;
    // The grid is sized from a count the host read some iterations ago; surplus workgroups
    // leave before touching the (device-scope, cross-XCD) tile counter.
    if (int(blockIdx.x) >= T) return;
    uint64_t* status = b.status[iter & 1];
    uint64_t* status_next = b.status[(iter + 1) & 1];
    const int32_t node_base = int32_t(2 * n - c);
    WindowLds w{s_lx, s_ly, s_lz, s_hx, s_hy, s_hz};

    for (;;) {
        if (threadIdx.x == 0) s_tile = int(atomicAdd(&b.tile_counter[iter & 1], 1u));
        __syncthreads();
        const int tile = s_tile;
        if (tile >= T) break;
        // the next iteration has at most ceil(c / BLOCK) tiles: clear ipt states per tile
        if (threadIdx.x < ipt) status_next[tile * ipt + threadIdx.x] = 0ull;
        if (threadIdx.x == 0 && tile == 0) b.tile_counter[(iter + 1) & 1] = 0u;
        const int a = tile * chunk;
        const int e = a + chunk < c ? a + chunk : c;
        const int w0 = a - 2 * R;
        const int wlen = (e - a) + 4 * R;
        for (int t = threadIdx.x; t < wlen; t += BLOCK) {
            int j = w0 + t;
            if (j >= 0 && j < c) {
                const float4* p = reinterpret_cast<const float4*>(in + j);
                float4 q0 = p[0], q1 = p[1];
                s_lx[t] = q0.x, s_ly[t] = q0.y, s_lz[t] = q0.z, s_hx[t] = q0.w;
                s_hy[t] = q1.x, s_hz[t] = q1.y, s_id[t] = __float_as_int(q1.z);
            }
        }
        __syncthreads();
        const int n0 = a - R < 0 ? 0 : a - R;
        const int n1 = e + R > c ? c : e + R;
        for (int i = n0 + threadIdx.x; i < n1; i += BLOCK) s_nn[i - (a - R)] = NearestNeighbour(w, i, w0, c, R);
        __syncthreads();

        // merge decisions for the tile's clusters (blocked: thread t owns kIpt in a row)
        uint32_t sflags = 0, mflags = 0;
        int32_t nnj[kIpt];
#pragma unroll
        for (int q = 0; q < kIpt; q++) {
            const int i = a + threadIdx.x * ipt + q;
            nnj[q] = -1;
            if (q < ipt && i < e) {
                const int j = s_nn[i - (a - R)];
                const bool mutual = s_nn[j - (a - R)] == i;
                nnj[q] = j;
                if (!(mutual && j < i)) sflags |= 1u << q;
                if (mutual && i < j) mflags |= 1u << q;
            }
        }
        const uint32_t mine = uint32_t(__popc(sflags)) | (uint32_t(__popc(mflags)) << 16);
        uint32_t total;
        const uint32_t before = BlockExclusiveScan<BLOCK>(mine, s_scan, total);
        const uint64_t aggS = total & 0xFFFFu, aggM = total >> 16;

        // decoupled look-back over earlier tiles (S in bits 31..61, M in bits 0..30)
        if (threadIdx.x < kWave) {
            uint64_t excl = 0;
            if (tile == 0) {
                if (threadIdx.x == 0)
                    __hip_atomic_store(&status[0], kFlagInc | (aggS << 31) | aggM, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            } else {
                if (threadIdx.x == 0)
                    __hip_atomic_store(&status[tile], kFlagAgg | (aggS << 31) | aggM, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                excl = WaveLookBack(status, tile);
                if (threadIdx.x == 0)
                    __hip_atomic_store(&status[tile], kFlagInc | (excl + ((aggS << 31) | aggM)), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
            if (threadIdx.x == 0) s_prefix = excl;
            if (threadIdx.x == 0 && tile == T - 1) {
                const int32_t next = int32_t(((excl >> 31) & kMask31) + aggS);
                b.ccount[iter + 1] = next;
                b.results->progress = next;
                b.results->multi_iters = iter + 1;
                b.results->flags = b.bounds[12];
            }
        }
        __syncthreads();
        const uint64_t excl = s_prefix;
        int32_t spos = int32_t((excl >> 31) & kMask31) + int32_t(before & 0xFFFFu);
        int32_t mpos = node_base + int32_t(excl & kMask31) + int32_t(before >> 16);
#pragma unroll
        for (int q = 0; q < kIpt; q++) {
            const int i = a + threadIdx.x * ipt + q;
            if (!(sflags & (1u << q))) continue;
            const int ii = i - w0;
            PNode o;
            if (mflags & (1u << q)) {
                const int jj = nnj[q] - w0;
                o.lo[0] = Min2(s_lx[ii], s_lx[jj]);
                o.lo[1] = Min2(s_ly[ii], s_ly[jj]);
                o.lo[2] = Min2(s_lz[ii], s_lz[jj]);
                o.hi[0] = Max2(s_hx[ii], s_hx[jj]);
                o.hi[1] = Max2(s_hy[ii], s_hy[jj]);
                o.hi[2] = Max2(s_hz[ii], s_hz[jj]);
                const int32_t id = mpos++;
                o.left = s_id[ii];
                o.right = s_id[jj];
                b.nodes[id] = o;
                b.parent[o.left] = id;
                b.parent[o.right] = id;
                StoreEval(b, id, InnerEval(o.lo, o.hi, LoadEval(b, o.left), LoadEval(b, o.right), max_leaf));
                if (b.ndirs > 3) UnionKdop(b, id, o.left, o.right);
                o.left = id;
                o.right = 0;
            } else {
                o.lo[0] = s_lx[ii], o.lo[1] = s_ly[ii], o.lo[2] = s_lz[ii];
                o.hi[0] = s_hx[ii], o.hi[1] = s_hy[ii], o.hi[2] = s_hz[ii];
                o.left = s_id[ii];
                o.right = 0;
            }
            out[spos++] = o;
        }
        __syncthreads();
    }
}

int PlocChunk(int block) { return block * kIpt; }
int64_t PlocStatusWords(int64_t n, int block) { return (n + block - 1) / block + kIpt; }
int PlocTail() { return kTail; }

hipError_t LaunchPlocIteration(const BuildBuffers& b, int iter, int radius, int block, int tail, int max_leaf, int grid,
                               int min_tiles, hipStream_t s) {
    if (grid < 1) grid = 1;
    if (min_tiles < 1) min_tiles = 1;
    switch (block) {
        case 128:
            hipLaunchKernelGGL(KPlocIteration<128>, dim3(grid), dim3(128), 0, s, b, iter, radius, tail, max_leaf,
                               min_tiles);
            break;
        case 512:
            hipLaunchKernelGGL(KPlocIteration<512>, dim3(grid), dim3(512), 0, s, b, iter, radius, tail, max_leaf,
                               min_tiles);
            break;
        default:
            hipLaunchKernelGGL(KPlocIteration<256>, dim3(grid), dim3(256), 0, s, b, iter, radius, tail, max_leaf,
                               min_tiles);
            break;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// tail: one workgroup runs every remaining iteration with the clusters (and their
// collapse data) resident in LDS
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(kTailThreads) void KPlocTail(BuildBuffers b, int iter, int R, int max_leaf) {
    constexpr int IPT = (kTail + kTailThreads - 1) / kTailThreads;
    __shared__ float s_lx[kTail], s_ly[kTail], s_lz[kTail], s_hx[kTail], s_hy[kTail], s_hz[kTail];
    __shared__ int32_t s_id[kTail];
    __shared__ int32_t s_nn[kTail];
    __shared__ int32_t s_cnt[kTail], s_ml[kTail], s_ht[kTail];
    __shared__ float s_cost[kTail];
    __shared__ uint32_t s_scan[kTailThreads / kWave + 1];
    const int64_t n = b.n;
    int c = b.ccount[iter];
    const PNode* in = b.clusters[iter & 1];
    for (int k = threadIdx.x; k < c; k += kTailThreads) {
        const float4* p = reinterpret_cast<const float4*>(in + k);
        float4 q0 = p[0], q1 = p[1];
        s_lx[k] = q0.x, s_ly[k] = q0.y, s_lz[k] = q0.z, s_hx[k] = q0.w;
        s_hy[k] = q1.x, s_hz[k] = q1.y, s_id[k] = __float_as_int(q1.z);
        const NodeEval e = LoadEval(b, s_id[k]);
        s_cnt[k] = e.count, s_cost[k] = e.cost, s_ml[k] = e.minleaf, s_ht[k] = e.height;
    }
    __syncthreads();
    WindowLds w{s_lx, s_ly, s_lz, s_hx, s_hy, s_hz};
    int it = iter;
    while (c > 1) {
        if (threadIdx.x == 0) b.ccount[it] = c;
        for (int i = threadIdx.x; i < c; i += kTailThreads) s_nn[i] = NearestNeighbour(w, i, 0, c, R);
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
                const int j = s_nn[i];
                const bool mutual = s_nn[j] == i;
                nnj[q] = j;
                if (!(mutual && j < i)) sflags |= 1u << q;
                if (mutual && i < j) mflags |= 1u << q;
            }
        }
        const uint32_t mine = uint32_t(__popc(sflags)) | (uint32_t(__popc(mflags)) << 16);
        uint32_t total;
        const uint32_t before = BlockExclusiveScan<kTailThreads>(mine, s_scan, total);
        if ((total >> 16) == 0u) {  // cannot happen for a symmetric distance; never spin on it
            if (threadIdx.x == 0) atomicOr(&b.bounds[12], 2u);
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
            const NodeEval ei{s_cnt[i], s_cost[i], s_ml[i], s_ht[i], 0};
            if (mflags & (1u << q)) {
                const int j = nnj[q];
                o.lo[0] = Min2(s_lx[i], s_lx[j]);
                o.lo[1] = Min2(s_ly[i], s_ly[j]);
                o.lo[2] = Min2(s_lz[i], s_lz[j]);
                o.hi[0] = Max2(s_hx[i], s_hx[j]);
                o.hi[1] = Max2(s_hy[i], s_hy[j]);
                o.hi[2] = Max2(s_hz[i], s_hz[j]);
                const int32_t id = mpos++;
                o.left = s_id[i];
                o.right = s_id[j];
                b.nodes[id] = o;
                b.parent[o.left] = id;
                b.parent[o.right] = id;
                const NodeEval ej{s_cnt[j], s_cost[j], s_ml[j], s_ht[j], 0};
                keep_e[q] = InnerEval(o.lo, o.hi, ei, ej, max_leaf);
                StoreEval(b, id, keep_e[q]);
                if (b.ndirs > 3) UnionKdop(b, id, o.left, o.right);
                o.left = id;
            } else {
                o.lo[0] = s_lx[i], o.lo[1] = s_ly[i], o.lo[2] = s_lz[i];
                o.hi[0] = s_hx[i], o.hi[1] = s_hy[i], o.hi[2] = s_hz[i];
                o.left = s_id[i];
                keep_e[q] = ei;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < IPT; q++) {
            if (!(sflags & (1u << q))) continue;
            const PNode& o = keep[q];
            const int k = spos++;
            s_lx[k] = o.lo[0], s_ly[k] = o.lo[1], s_lz[k] = o.lo[2];
            s_hx[k] = o.hi[0], s_hy[k] = o.hi[1], s_hz[k] = o.hi[2];
            s_id[k] = o.left;
            s_cnt[k] = keep_e[q].count, s_cost[k] = keep_e[q].cost, s_ml[k] = keep_e[q].minleaf;
            s_ht[k] = keep_e[q].height;
        }
        c = int(total & 0xFFFFu);
        it++;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        b.ccount[it] = c;
        b.iter_done[0] = it;
        b.results->iter_done = it;
        b.results->tail_iters = it - iter;
    }
}

hipError_t LaunchPlocTail(const BuildBuffers& b, int iter, int radius, int max_leaf, hipStream_t s) {
    hipLaunchKernelGGL(KPlocTail, dim3(1), dim3(kTailThreads), 0, s, b, iter, radius, max_leaf);
    return hipGetLastError();
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

__global__ __launch_bounds__(256) void KClassify(BuildBuffers b, int max_leaf) {
    const int64_t n = b.n, nn = 2 * n - 1;
    int kept = 0, leaves = 0;
    for (int64_t id = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; id < nn; id += int64_t(gridDim.x) * blockDim.x) {
        const bool in = Inside(b, int32_t(id), max_leaf);
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
    // block-reduce the two counters
    __shared__ int s_k, s_l;
    if (threadIdx.x == 0) s_k = s_l = 0;
    __syncthreads();
    if (kept) atomicAdd(&s_k, kept);
    if (leaves) atomicAdd(&s_l, leaves);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_k) atomicAdd(&b.counters[0], s_k);
        if (s_l) atomicAdd(&b.counters[1], s_l);
    }
}

// Gathers the primitives of every final leaf, and writes the node of every kept inner node.
// keptflag has been scanned in place into exclusive ranks and leafslot into begin offsets.
__global__ __launch_bounds__(256) void KEmit(BuildBuffers b, int max_leaf) {
    const int64_t n = b.n, nn = 2 * n - 1;
    const int32_t K = b.counters[0];
    const int32_t* __restrict__ rank = b.keptflag;   // exclusive ranks
    const int32_t* __restrict__ begin = b.leafslot;  // exclusive offsets by minleaf
    auto kept = [&](int32_t id) -> bool {
        if (id < n) return false;
        // kept iff its rank differs from the next one's (or it is the last with a 1)
        return !b.collapsed[id] && !Inside(b, id, max_leaf);
    };
    auto leaf_ref = [&](int32_t id) -> int32_t { return ~((begin[b.minleaf[id]] << 4) | b.count[id]); };
    __shared__ int4 s_tri[kMaxTriples];
    if (b.ndirs > 0) {
        for (int t = threadIdx.x; t < b.ntriples; t += blockDim.x) s_tri[t] = b.triples[t];
        __syncthreads();
    }
    const float pad = b.ndirs > 0 ? KdopPad(b.bounds) : 0.0f;
    for (int64_t id = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; id < nn; id += int64_t(gridDim.x) * blockDim.x) {
        const bool col = b.collapsed[id] != 0;
        const bool in = Inside(b, int32_t(id), max_leaf);
        if (col && !in) {
            // gather this final leaf's primitives: left-first DFS of its PLOC subtree
            int32_t pos = begin[b.minleaf[id]];
            int32_t st[kMaxLeaf + 1];
            int sp = 0;
            st[sp++] = int32_t(id);
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
                f.left = leaf_ref(int32_t(id));
                f.right = ~0;
                f.pad0 = f.pad1 = 0;
                if (b.ndirs > 0) {
                    float g[6];
                    const uint32_t d = FitSobb(b, s_tri, int32_t(id), pad, g);
                    float* fl = &f.lx0;
                    for (int k = 0; k < 6; k++) fl[k] = fl[6 + k] = g[k];
                    f.pad0 = int32_t(d | (d << 12));
                }
                b.fnodes[0] = f;
            }
        } else if (id >= n && !col && !in) {
            const PNode nd = b.nodes[id];
            const int32_t left = kept(nd.left) ? K - 1 - rank[nd.left - n] : leaf_ref(nd.left);
            const int32_t right = kept(nd.right) ? K - 1 - rank[nd.right - n] : leaf_ref(nd.right);
            FNode* out = b.fnodes + (K - 1 - rank[id - n]);
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
}

// SAH of the final tree in FP64 (Ct = Ci = 1, divided by SA(root) on the host); the node
// count is the number of kept inner nodes (0 when the whole scene is one leaf).
// SOBB area 2 (wa wb + wb wc + wa wc) / |det(n_a, n_b, n_c)| in FP64.
__device__ double SobbAreaF64(const float* f, uint32_t d, const float* dirs) {
    const float* n1 = dirs + 3 * (d & 15);
    const float* n2 = dirs + 3 * ((d >> 4) & 15);
    const float* n3 = dirs + 3 * ((d >> 8) & 15);
    const double det = double(n1[0]) * (double(n2[1]) * n3[2] - double(n2[2]) * n3[1]) +
                       double(n1[1]) * (double(n2[2]) * n3[0] - double(n2[0]) * n3[2]) +
                       double(n1[2]) * (double(n2[0]) * n3[1] - double(n2[1]) * n3[0]);
    const double wa = double(f[1]) - f[0], wb = double(f[3]) - f[2], wc = double(f[5]) - f[4];
    return 2.0 * (wa * wb + wb * wc + wa * wc) / fabs(det);
}

__global__ __launch_bounds__(256) void KSah(const FNode* nodes, const int32_t* counters, double* out,
                                            const float* dirs) {
    const int32_t count = counters[0];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const FNode f = nodes[i];
        auto sa = [](float x0, float x1, float y0, float y1, float z0, float z1) {
            double sx = double(x1) - x0, sy = double(y1) - y0, sz = double(z1) - z0;
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
        acc += f.left >= 0 ? sl : sl * double((~f.left) & 15);
        acc += f.right >= 0 ? sr : sr * double((~f.right) & 15);
    }
    __shared__ double s[256];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if (int(threadIdx.x) < d) s[threadIdx.x] += s[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0 && s[0] != 0.0) atomicAdd(out, s[0]);
}

hipError_t LaunchLayout(const BuildBuffers& b, int max_leaf, hipStream_t s) {
    const int64_t nn = 2 * b.n - 1;
    int grid = int(std::min<int64_t>(4096, (nn + 255) / 256));
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(KClassify, dim3(grid), dim3(256), 0, s, b, max_leaf);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = LaunchExclusiveScan(b.leafslot, b.leafslot, b.n, b.scan_tmp, b.counters + 4, s);
    if (e != hipSuccess) return e;
    if (b.n > 1) {
        e = LaunchExclusiveScan(b.keptflag, b.keptflag, b.n - 1, b.scan_tmp, b.counters + 5, s);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(KEmit, dim3(grid), dim3(256), 0, s, b, max_leaf);
    return hipGetLastError();
}

// Copies the build's small results into host-mapped memory (one thread, after the SAH sum).
__global__ void KFinish(BuildBuffers b) {
    BuildResults* r = b.results;
    const int64_t root = 2 * b.n - 2;
    r->kept = b.counters[0];
    r->leaves = b.counters[1];
    r->root_height = b.height[root];
    r->sah = b.sah[0];
    r->flags = b.bounds[12];
    const PNode nd = b.nodes[root];
    for (int a = 0; a < 3; a++) {
        r->root_lo[a] = nd.lo[a];
        r->root_hi[a] = nd.hi[a];
    }
}

hipError_t LaunchFinish(const BuildBuffers& b, hipStream_t s) {
    hipLaunchKernelGGL(KFinish, dim3(1), dim3(1), 0, s, b);
    return hipGetLastError();
}

hipError_t LaunchSah(const FNode* nodes, const int32_t* counters, double* out, const float* sobb_dirs,
                     hipStream_t s) {
    hipLaunchKernelGGL(KSah, dim3(512), dim3(256), 0, s, nodes, counters, out, sobb_dirs);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// exclusive scan of int32 (in place allowed): block sums, one-workgroup scan, downsweep
// ---------------------------------------------------------------------------------------
namespace {
constexpr int kScanBlock = 256;
constexpr int kScanIpt = 8;
constexpr int kScanTile = kScanBlock * kScanIpt;
}  // namespace

size_t ScanTempInts(int64_t n) { return size_t((n + kScanTile - 1) / kScanTile) + 1; }

__global__ __launch_bounds__(kScanBlock) void KScanReduce(const int32_t* in, int64_t n, int32_t* sums) {
    const int64_t base = int64_t(blockIdx.x) * kScanTile;
    int32_t acc = 0;
    for (int k = threadIdx.x; k < kScanTile; k += kScanBlock) {
        int64_t i = base + k;
        if (i < n) acc += in[i];
    }
    __shared__ int32_t s[kScanBlock];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int d = kScanBlock / 2; d > 0; d >>= 1) {
        if (int(threadIdx.x) < d) s[threadIdx.x] += s[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(1024) void KScanSums(int32_t* sums, int64_t m, int32_t* total) {
    __shared__ uint32_t s_scan[1024 / kWave + 1];
    __shared__ int32_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < m; base += 1024) {
        int64_t i = base + threadIdx.x;
        int32_t v = i < m ? sums[i] : 0;
        uint32_t tot;
        uint32_t ex = BlockExclusiveScan<1024>(uint32_t(v), s_scan, tot);
        const int32_t carry = s_carry;
        if (i < m) sums[i] = carry + int32_t(ex);
        __syncthreads();
        if (threadIdx.x == 0) s_carry = carry + int32_t(tot);
        __syncthreads();
    }
    if (threadIdx.x == 0) total[0] = s_carry;
}

__global__ __launch_bounds__(kScanBlock) void KScanDown(const int32_t* in, int32_t* out, int64_t n,
                                                        const int32_t* sums) {
    __shared__ uint32_t s_scan[kScanBlock / kWave + 1];
    const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanIpt;
    int32_t v[kScanIpt];
    int32_t acc = 0;
#pragma unroll
    for (int q = 0; q < kScanIpt; q++) {
        int64_t i = base + q;
        v[q] = i < n ? in[i] : 0;
        acc += v[q];
    }
    uint32_t tot;
    int32_t run = sums[blockIdx.x] + int32_t(BlockExclusiveScan<kScanBlock>(uint32_t(acc), s_scan, tot));
#pragma unroll
    for (int q = 0; q < kScanIpt; q++) {
        int64_t i = base + q;
        if (i < n) out[i] = run;
        run += v[q];
    }
}

hipError_t LaunchExclusiveScan(const int32_t* in, int32_t* out, int64_t n, int32_t* tmp, int32_t* total,
                               hipStream_t s) {
    const int64_t blocks = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(KScanReduce, dim3(unsigned(blocks)), dim3(kScanBlock), 0, s, in, n, tmp);
    hipLaunchKernelGGL(KScanSums, dim3(1), dim3(1024), 0, s, tmp, blocks, total);
    hipLaunchKernelGGL(KScanDown, dim3(unsigned(blocks)), dim3(kScanBlock), 0, s, in, out, n, tmp);
    return hipGetLastError();
}

}  // namespace dpl
