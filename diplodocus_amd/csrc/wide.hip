// Wide trees for gfx950: the two-child final tree converted into 4- or 8-child nodes whose
// child boxes are 8-bit offsets from the node's minimum corner in power-of-two units per
// axis (64 B nodes for 4 children, 96 B for 8), and the persistent trace kernel over them.
//
// Conversion is one thread per node of the two-child tree: wide node i is node i opened
// greedily (the inner child of largest surface area first, ties to the leftmost) until it
// has W children, so every node is converted independently and the layout is the two-child
// tree's (only the nodes a traversal from node 0 reaches are ever read). Quantization keeps
// every decoded box a superset of the exact one: the exponent is the smallest with
// fmaf(255, 2^e, lo) >= hi, the low bound the largest q with fmaf(q, 2^e, lo) <= lo_k, the
// high bound the smallest q with fmaf(q, 2^e, lo) >= hi_k. The traversal decodes with the
// same fmaf, so the slab test sees a box that contains the child and the closest hit is the
// two-child trace's. SOBB trees are first refitted to axis-aligned boxes, bottom-up one tree
// level per launch.
#include <hip/hip_runtime.h>

#include "dpl_internal.h"
#include "traverse.h"

namespace dpl {

namespace {
constexpr int kWideBlock = 256;
constexpr int kWideStackLds = 16;
constexpr int kWideWgPerCu = 8;
constexpr int kWidePending = 4;  // leaves a lane keeps pending (as the two-child trace)

__device__ __forceinline__ bool EmptyLeaf(int32_t r) { return r < 0 && ((~r) & 15) == 0; }
__device__ __forceinline__ float ScaleOf(uint32_t b) { return __uint_as_float(b << 23); }

// Child box of a final node: side 0 floats 0-3, 8-9; side 1 floats 4-7, 10-11.
__device__ __forceinline__ void ChildBox(const FNode& f, int side, float lo[3], float hi[3]) {
    if (side == 0) {
        lo[0] = f.lx0, hi[0] = f.lx1, lo[1] = f.ly0, hi[1] = f.ly1, lo[2] = f.lz0, hi[2] = f.lz1;
    } else {
        lo[0] = f.rx0, hi[0] = f.rx1, lo[1] = f.ry0, hi[1] = f.ry1, lo[2] = f.rz0, hi[2] = f.rz1;
    }
}

__device__ __forceinline__ void StoreChildBox(FNode* f, int side, const float lo[3], const float hi[3]) {
    float* p = reinterpret_cast<float*>(f);
    const int o = side == 0 ? 0 : 4, oz = side == 0 ? 8 : 10;
    p[o] = lo[0], p[o + 1] = hi[0], p[o + 2] = lo[1], p[o + 3] = hi[1];
    p[oz] = lo[2], p[oz + 1] = hi[2];
}

// Smallest biased exponent b in [1, 254] with fmaf(255, 2^(b-127), L) >= H (exists for finite
// bounds: b = 254 overflows to +inf).
__device__ uint32_t ChooseExponent(float L, float H) {
    const double ext = double(H) - double(L);
    int b = 1;
    if (ext > 0.0) {
        int e = 0;
        (void)frexp(ext / 255.0, &e);
        b = min(254, max(1, e + 127));
    }
    while (b > 1 && fmaf(255.0f, ScaleOf(uint32_t(b - 1)), L) >= H) b--;
    while (b < 254 && !(fmaf(255.0f, ScaleOf(uint32_t(b)), L) >= H)) b++;
    return uint32_t(b);
}

__device__ uint32_t QuantLo(float lo, float L, float s) {
    const double est = floor((double(lo) - double(L)) / double(s));
    int q = int(fmin(255.0, fmax(0.0, est)));
    while (q < 255 && fmaf(float(q + 1), s, L) <= lo) q++;
    while (q > 0 && fmaf(float(q), s, L) > lo) q--;
    return uint32_t(q);
}

__device__ uint32_t QuantHi(float hi, float L, float s) {
    const double est = ceil((double(hi) - double(L)) / double(s));
    int q = int(fmin(255.0, fmax(0.0, est)));
    while (q > 0 && fmaf(float(q - 1), s, L) >= hi) q--;
    while (q < 255 && !(fmaf(float(q), s, L) >= hi)) q++;
    return uint32_t(q);
}
}  // namespace

// ---- AABB refit of a final tree --------------------------------------------------------
// Level by level, one launch per level: launch 1 boxes the leaf children of every node and
// completes the nodes whose children are all leaves; launch L completes the nodes whose
// inner children were completed by earlier launches (done[c] in [1, L)). A node of height h
// completes in launch h + 1, so height + 1 launches finish the root. Values written by
// earlier launches are visible at a kernel boundary; one written in the current launch is
// never read (done[c] == L counts as not ready), so no fences or atomics are needed.

__device__ __forceinline__ void UnionOf(const FNode& f, float* box6) {
    box6[0] = Min2(f.lx0, f.rx0), box6[1] = Min2(f.ly0, f.ry0), box6[2] = Min2(f.lz0, f.rz0);
    box6[3] = Max2(f.lx1, f.rx1), box6[4] = Max2(f.ly1, f.ry1), box6[5] = Max2(f.lz1, f.rz1);
}

__global__ __launch_bounds__(kWideBlock) void KRefitLeaves(const FNode* __restrict__ nodes, int32_t K,
                                                           const float4* __restrict__ tris, int32_t nprim,
                                                           FNode* __restrict__ out, float* __restrict__ box,
                                                           int32_t* __restrict__ done) {
    const int32_t s = int32_t(blockIdx.x) * kWideBlock + int32_t(threadIdx.x);
    if (s >= K) return;
    const int4 d = reinterpret_cast<const int4*>(nodes + s)[3];
    FNode o;
    float* ob = reinterpret_cast<float*>(&o);
    for (int k = 0; k < 12; k++) ob[k] = 0.0f;
    o.left = d.x, o.right = d.y, o.pad0 = 0, o.pad1 = 0;
    float lo[3], hi[3];
    for (int side = 0; side < 2; side++) {
        const int32_t ref = side == 0 ? d.x : d.y;
        if (ref >= 0 || EmptyLeaf(ref)) continue;  // inner: a later launch; empty: repeats the left box
        const int32_t packed = ~ref, b0 = packed >> 4;
        const int32_t cnt = int64_t(b0) + (packed & 15) <= nprim ? (packed & 15) : 0;
        for (int a = 0; a < 3; a++) lo[a] = 3.40282347e38f, hi[a] = -3.40282347e38f;
        for (int32_t k = b0; k < b0 + cnt; k++) {
            const float4 v[3] = {tris[3 * int64_t(k)], tris[3 * int64_t(k) + 1], tris[3 * int64_t(k) + 2]};
            for (int j = 0; j < 3; j++) {
                lo[0] = Min2(lo[0], v[j].x), hi[0] = Max2(hi[0], v[j].x);
                lo[1] = Min2(lo[1], v[j].y), hi[1] = Max2(hi[1], v[j].y);
                lo[2] = Min2(lo[2], v[j].z), hi[2] = Max2(hi[2], v[j].z);
            }
        }
        StoreChildBox(&o, side, lo, hi);
    }
    if (EmptyLeaf(d.y)) {
        float l2[3], h2[3];
        ChildBox(o, 0, l2, h2);
        StoreChildBox(&o, 1, l2, h2);
    }
    const bool complete = d.x < 0 && d.y < 0;
    out[s] = o;
    if (complete) UnionOf(o, box + 6 * int64_t(s));
    done[s] = complete ? 1 : 0;
}

__global__ __launch_bounds__(kWideBlock) void KRefitLevel(const FNode* __restrict__ nodes, int32_t K, FNode* out,
                                                          float* box, int32_t* done, int32_t level) {
    const int32_t s = int32_t(blockIdx.x) * kWideBlock + int32_t(threadIdx.x);
    if (s >= K || done[s] != 0) return;
    const int4 d = reinterpret_cast<const int4*>(nodes + s)[3];
    for (int side = 0; side < 2; side++) {
        const int32_t c = side == 0 ? d.x : d.y;
        if (c < 0) continue;
        if (c >= K) return;  // not a valid tree: never completes
        const int32_t dc = done[c];
        if (dc == 0 || dc >= level) return;
    }
    FNode f = out[s];
    for (int side = 0; side < 2; side++) {
        const int32_t c = side == 0 ? d.x : d.y;
        if (c < 0) continue;
        const float* cb = box + 6 * int64_t(c);
        const float lo[3] = {cb[0], cb[1], cb[2]}, hi[3] = {cb[3], cb[4], cb[5]};
        StoreChildBox(&f, side, lo, hi);
    }
    out[s] = f;
    UnionOf(f, box + 6 * int64_t(s));
    done[s] = level;
}

hipError_t LaunchRefitAabb(const FNode* nodes, int32_t K, const float4* tris, int32_t nprim, int32_t height, FNode* out,
                           float* box, int32_t* done, hipStream_t s) {
    if (K <= 0) return hipSuccess;
    const int grid = (K + kWideBlock - 1) / kWideBlock;
    hipLaunchKernelGGL(KRefitLeaves, dim3(grid), dim3(kWideBlock), 0, s, nodes, K, tris, nprim, out, box, done);
    for (int32_t level = 2; level <= height + 1; level++)
        hipLaunchKernelGGL(KRefitLevel, dim3(grid), dim3(kWideBlock), 0, s, nodes, K, out, box, done, level);
    return hipGetLastError();
}

// ---- conversion ------------------------------------------------------------------------

// mt (batched BVHs, else null): node slot i belongs to the mesh whose slots start at the
// largest node_off <= i; that mesh's references are local to its slots. Slots past a mesh's
// node count are never reached and get a zero node.
template <int W>
__global__ __launch_bounds__(kWideBlock) void KWiden(const FNode* __restrict__ all_nodes, int32_t K_all,
                                                     float4* __restrict__ out, const BatchTree* __restrict__ mt,
                                                     int32_t mesh_count) {
    const int32_t gi = int32_t(blockIdx.x) * kWideBlock + int32_t(threadIdx.x);
    if (gi >= K_all) return;
    const FNode* nodes = all_nodes;
    int32_t K = K_all, i = gi;
    if (mt) {
        int lo = 0, hi = mesh_count - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (mt[mid].node_off <= gi) lo = mid;
            else hi = mid - 1;
        }
        const BatchTree bt = mt[lo];
        i = int32_t(gi - bt.node_off);
        if (i >= bt.node_count) {
            float4* o = out + int64_t(gi) * (WideNodeBytes(W) / 16);
            for (int j = 0; j < WideNodeBytes(W) / 16; j++) o[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            return;
        }
        nodes = all_nodes + bt.node_off;
        K = bt.node_count;
    }
    int32_t ref[W];
    float lo[W][3], hi[W][3];
    double area[W];
    int c = 0;
    auto half_area = [](const float* l, const float* h) {
        const double dx = double(h[0]) - double(l[0]);
        const double dy = double(h[1]) - double(l[1]);
        const double dz = double(h[2]) - double(l[2]);
        return dx * dy + dy * dz + dz * dx;
    };
#pragma unroll
    for (int k = 0; k < W; k++) {
        ref[k] = ~0, area[k] = 0.0;
        for (int a = 0; a < 3; a++) lo[k][a] = hi[k][a] = 0.0f;
    }
    {
        const FNode f = nodes[i];
        for (int side = 0; side < 2; side++) {
            const int32_t r = side == 0 ? f.left : f.right;
            if (EmptyLeaf(r)) continue;
            float l[3], h[3];
            ChildBox(f, side, l, h);
#pragma unroll
            for (int k = 0; k < 2; k++)
                if (k == c) {
                    ref[k] = r;
                    for (int a = 0; a < 3; a++) lo[k][a] = l[a], hi[k][a] = h[a];
                    area[k] = half_area(l, h);
                }
            c++;
        }
    }
    while (c < W) {
        int best = -1;
        double ba = 0.0;
#pragma unroll
        for (int p = 0; p < W; p++)
            if (p < c && ref[p] >= 0 && ref[p] < K && (best < 0 || area[p] > ba)) best = p, ba = area[p];
        if (best < 0) break;
        int32_t br = 0;
#pragma unroll
        for (int p = 0; p < W; p++)
            if (p == best) br = ref[p];
        const FNode g = nodes[br];
        float l0[3], h0[3], l1[3], h1[3];
        ChildBox(g, 0, l0, h0);
        ChildBox(g, 1, l1, h1);
        const double a0 = half_area(l0, h0), a1 = half_area(l1, h1);
        // shift entries best+1 .. c-1 one to the right, then place the two children
#pragma unroll
        for (int p = W - 1; p >= 1; p--) {
            if (p > best + 1 && p <= c) {
                ref[p] = ref[p - 1], area[p] = area[p - 1];
                for (int a = 0; a < 3; a++) lo[p][a] = lo[p - 1][a], hi[p][a] = hi[p - 1][a];
            }
        }
#pragma unroll
        for (int p = 0; p < W; p++) {
            if (p == best) {
                ref[p] = g.left, area[p] = a0;
                for (int a = 0; a < 3; a++) lo[p][a] = l0[a], hi[p][a] = h0[a];
            } else if (p == best + 1) {
                ref[p] = g.right, area[p] = a1;
                for (int a = 0; a < 3; a++) lo[p][a] = l1[a], hi[p][a] = h1[a];
            }
        }
        c++;
    }
    float origin[3];
    uint32_t bexp[3];
    uint32_t planes[6][W / 4];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float L = lo[0][a], H = hi[0][a];
#pragma unroll
        for (int p = 1; p < W; p++)
            if (p < c) L = Min2(L, lo[p][a]), H = Max2(H, hi[p][a]);
        origin[a] = c > 0 ? L : 0.0f;
        bexp[a] = c > 0 ? ChooseExponent(L, H) : 1u;
        const float sc = ScaleOf(bexp[a]);
#pragma unroll
        for (int j = 0; j < W / 4; j++) planes[2 * a][j] = planes[2 * a + 1][j] = 0u;
#pragma unroll
        for (int p = 0; p < W; p++) {
            if (p < c) {
                planes[2 * a][p / 4] |= QuantLo(lo[p][a], L, sc) << (8 * (p % 4));
                planes[2 * a + 1][p / 4] |= QuantHi(hi[p][a], L, sc) << (8 * (p % 4));
            }
        }
    }
    const uint32_t meta = bexp[0] | (bexp[1] << 8) | (bexp[2] << 16) | (uint32_t(c) << 24);
    float4* o = out + int64_t(gi) * (WideNodeBytes(W) / 16);
    o[0] = make_float4(origin[0], origin[1], origin[2], __uint_as_float(meta));
#pragma unroll
    for (int j = 0; j < W / 4; j++)
        o[1 + j] = make_float4(__int_as_float(ref[4 * j]), __int_as_float(ref[4 * j + 1]),
                               __int_as_float(ref[4 * j + 2]), __int_as_float(ref[4 * j + 3]));
    uint32_t words[WideNodeBytes(W) / 4 - 4 - W] = {};
#pragma unroll
    for (int pl = 0; pl < 6; pl++)
#pragma unroll
        for (int j = 0; j < W / 4; j++) words[pl * (W / 4) + j] = planes[pl][j];
#pragma unroll
    for (int j = 0; j < (WideNodeBytes(W) / 4 - 4 - W) / 4; j++)
        o[1 + W / 4 + j] = make_float4(__uint_as_float(words[4 * j]), __uint_as_float(words[4 * j + 1]),
                                       __uint_as_float(words[4 * j + 2]), __uint_as_float(words[4 * j + 3]));
}

hipError_t LaunchWiden(const FNode* nodes, int32_t K, int width, void* out, const BatchTree* mt, int32_t mesh_count,
                       hipStream_t s) {
    if (K <= 0) return hipSuccess;
    const int grid = (K + kWideBlock - 1) / kWideBlock;
    float4* o = static_cast<float4*>(out);
    if (width == 4)
        hipLaunchKernelGGL(KWiden<4>, dim3(grid), dim3(kWideBlock), 0, s, nodes, K, o, mt, mesh_count);
    else if (width == 8)
        hipLaunchKernelGGL(KWiden<8>, dim3(grid), dim3(kWideBlock), 0, s, nodes, K, o, mt, mesh_count);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// ---- traversal -------------------------------------------------------------------------

// Compare-exchange on (tnear, slot): a strict total order, so any sorting network gives the
// one order (tnear ascending, ties to the lower slot; misses carry slot + W and +inf).
__device__ __forceinline__ void Cas(float& ta, int& ka, int32_t& ra, float& tb, int& kb, int32_t& rb) {
    const bool sw = (ta > tb) | ((ta == tb) & (ka > kb));
    const float t = sw ? tb : ta;
    tb = sw ? ta : tb;
    ta = t;
    const int k = sw ? kb : ka;
    kb = sw ? ka : kb;
    ka = k;
    const int32_t r = sw ? rb : ra;
    rb = sw ? ra : rb;
    ra = r;
}

template <int W>
__device__ __forceinline__ void SortChildren(float* t, int* k, int32_t* r) {
#define DPL_CAS(a, b) Cas(t[a], k[a], r[a], t[b], k[b], r[b])
    if (W == 4) {
        DPL_CAS(0, 1), DPL_CAS(2, 3), DPL_CAS(0, 2), DPL_CAS(1, 3), DPL_CAS(1, 2);
    } else {
        DPL_CAS(0, 1), DPL_CAS(2, 3), DPL_CAS(4, 5), DPL_CAS(6, 7);
        DPL_CAS(0, 2), DPL_CAS(1, 3), DPL_CAS(4, 6), DPL_CAS(5, 7);
        DPL_CAS(1, 2), DPL_CAS(5, 6);
        DPL_CAS(0, 4), DPL_CAS(1, 5), DPL_CAS(2, 6), DPL_CAS(3, 7);
        DPL_CAS(2, 4), DPL_CAS(3, 5);
        DPL_CAS(1, 2), DPL_CAS(3, 4), DPL_CAS(5, 6);
    }
#undef DPL_CAS
}

// One wide node of an ordered traversal: slab tests of all W children without a branch (a
// wave's lanes sit at nodes with different child counts), the hit children sorted by
// (tnear, slot), all but the nearest pushed farthest first, ref set to the nearest. False when
// no child is hit (the caller pops the next entry).
template <int W, class Stack>
__device__ __forceinline__ bool WideNodeStep(const WideTree& tree, const WtRay& w, float best_t, int32_t& ref,
                                             Stack& st) {
    constexpr int NB = WideNodeBytes(W) / 16;  // float4 per node
    const float4* np = tree.nodes + int64_t(ref) * NB;
    uint32_t wd[NB * 4];
#pragma unroll
    for (int j = 0; j < NB; j++) {
        const uint4 v = reinterpret_cast<const uint4*>(np)[j];
        wd[4 * j] = v.x, wd[4 * j + 1] = v.y, wd[4 * j + 2] = v.z, wd[4 * j + 3] = v.w;
    }
    const float ox = __uint_as_float(wd[0]), oy = __uint_as_float(wd[1]), oz = __uint_as_float(wd[2]);
    const uint32_t meta = wd[3];
    const float sx = ScaleOf(meta & 255u), sy = ScaleOf((meta >> 8) & 255u), sz = ScaleOf((meta >> 16) & 255u);
    const int c = int(meta >> 24);
    float key[W];
    int slot[W];
    int32_t val[W];
    int nh = 0;
#pragma unroll
    for (int k = 0; k < W; k++) {
        constexpr int Q = 4 + W;  // first plane word
        const int sh = 8 * (k % 4), wj = k / 4;
        auto q = [&](int pl) { return float((wd[Q + pl * (W / 4) + wj] >> sh) & 255u); };
        const float x0 = fmaf(q(0), sx, ox), x1 = fmaf(q(1), sx, ox);
        const float y0 = fmaf(q(2), sy, oy), y1 = fmaf(q(3), sy, oy);
        const float z0 = fmaf(q(4), sz, oz), z1 = fmaf(q(5), sz, oz);
        float tn;
        const bool hit = Box(w, best_t, x0, x1, y0, y1, z0, z1, tn) & (k < c);
        key[k] = hit ? tn : __builtin_huge_valf();
        slot[k] = hit ? k : k + W;
        val[k] = int32_t(wd[4 + k]);
        nh += hit ? 1 : 0;
    }
    if (nh == 0) return false;
    SortChildren<W>(key, slot, val);
#pragma unroll
    for (int j = W - 1; j >= 1; j--)
        if (j < nh) st.Push(val[j]);
    ref = val[0];
    return true;
}

// Ordered traversal of a wide tree: the hit children sorted by (tnear, slot), all but the
// nearest pushed farthest first; leaves are pushed like nodes. Every child's slab test runs
// without a branch (a wave's lanes sit at nodes with different child counts). ANY stops at
// the first hit.
template <bool ANY, int W, class Stack>
__device__ __forceinline__ HitResult TraverseWide(const WideTree& tree, const WtRay& w, bool cull, Stack& st,
                                                  TraceCounters& cnt) {
    HitResult r{0.0f, 0.0f, 0.0f, -1, -1};
    float best_t = w.tmax;
    st.sp = 0;
    if (tree.node_count <= 0) return r;
    int32_t ref = 0;
    int64_t budget = 2 * int64_t(tree.node_count) + 2;
    for (;;) {
        if (--budget < 0 || ref >= tree.node_count) break;
        if (ref >= 0) {
            cnt.nodes++;
            if (!WideNodeStep<W>(tree, w, best_t, ref, st)) {
                if (st.sp == 0) break;
                ref = st.Pop();
            }
        } else {
            if (TestLeaf<ANY>(tree.tris, tree.prim_count, ref, w, cull, best_t, r, cnt) || st.sp == 0) break;
            ref = st.Pop();
        }
    }
    return r;
}

// BATCH: ray i is traced against mesh ray_mesh[i] (a mesh id out of range gives a miss);
// tree.nodes / tree.tris are then the bases of all meshes' wide nodes and triangles.
template <int W, bool COUNT, bool BATCH>
__global__ __launch_bounds__(kWideBlock) void KTraceWide(WideTree tree, const float4* __restrict__ rays, int64_t n,
                                                         uint32_t flags, dpl_hit* __restrict__ hits,
                                                         uint32_t* counter, int32_t* spill,
                                                         unsigned long long* counters,
                                                         const int32_t* __restrict__ ray_mesh,
                                                         const BatchTree* __restrict__ mt, int32_t mesh_count) {
    __shared__ int32_t s_stack[kWideStackLds][kWideBlock];
    __shared__ uint32_t s_base[kWideBlock / kWave];
    const bool any = flags & DPL_TRACE_ANY, cull = flags & DPL_TRACE_CULL_BACK;
    const int lane = LaneId(), wid = threadIdx.x / kWave;
    LaneStackInterleaved<kWideStackLds, kWideBlock> st{&s_stack[0][threadIdx.x],
                                                       spill + int64_t(blockIdx.x) * kWideBlock + threadIdx.x,
                                                       int64_t(gridDim.x) * kWideBlock, 0};
    TraceCounters cnt{0u, 0u};
    for (;;) {
        if (lane == 0) s_base[wid] = atomicAdd(counter, uint32_t(kWave));
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int64_t base = s_base[wid];
        if (base >= n) break;
        const int64_t i = base + lane;
        if (i < n) {
            const float4 r0 = rays[2 * i], r1 = rays[2 * i + 1];
            const WtRay w = PrepareRay(r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w);
            WideTree t = tree;
            bool valid = true;
            if (BATCH) {
                const int32_t m = ray_mesh[i];
                valid = m >= 0 && m < mesh_count;
                if (valid) {
                    const BatchTree bt = mt[m];
                    t.nodes = tree.nodes + bt.node_off * (WideNodeBytes(W) / 16);
                    t.tris = tree.tris + 3 * bt.tri_off;
                    t.node_count = bt.node_count;
                    t.prim_count = bt.prim_count;
                }
            }
            HitResult h{0.0f, 0.0f, 0.0f, -1, -1};
            if (valid)
                h = any ? TraverseWide<true, W>(t, w, cull, st, cnt) : TraverseWide<false, W>(t, w, cull, st, cnt);
            dpl_hit o;
            o.t = h.t;
            o.b1 = h.b1;
            o.b2 = h.b2;
            o.prim = h.prim;
            hits[i] = o;
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (COUNT) {
        unsigned long long a = cnt.nodes, b = cnt.tris;
#pragma unroll
        for (int d = kWave / 2; d > 0; d >>= 1) {
            a += __shfl_xor(a, d, kWave);
            b += __shfl_xor(b, d, kWave);
        }
        if (lane == 0) {
            atomicAdd(&counters[0], a);
            atomicAdd(&counters[1], b);
        }
    }
}

// The wide trace with per-lane ray replacement and up to Q pending leaves per lane, tested one
// triangle per step (as the two-child refill kernel, trace.hip): same closest hits.
template <int W, bool BATCH, bool ANY, int Q>
__global__ __launch_bounds__(kWideBlock) void KTraceWideRefill(WideTree tree, const float4* __restrict__ rays,
                                                               int64_t n, bool cull, dpl_hit* __restrict__ hits,
                                                               uint32_t* counter, int32_t* spill,
                                                               const int32_t* __restrict__ ray_mesh,
                                                               const BatchTree* __restrict__ mt, int32_t mesh_count,
                                                               int refill, const int32_t* __restrict__ order,
                                                               int shares, int64_t share_len) {
    constexpr int32_t kEnd = 0x7FFFFFFF;
    __shared__ int32_t s_stack[kWideStackLds][kWideBlock];
    __shared__ uint32_t s_base[kWideBlock / kWave];
    const int lane = LaneId(), wid = threadIdx.x / kWave;
    const uint64_t below = (1ull << lane) - 1ull;
    LaneStackInterleaved<kWideStackLds, kWideBlock> st{&s_stack[0][threadIdx.x],
                                                       spill + int64_t(blockIdx.x) * kWideBlock + threadIdx.x,
                                                       int64_t(gridDim.x) * kWideBlock, 0};
    int64_t ray = -1;
    WtRay w{};
    WideTree t = tree;
    float best_t = 0.0f;
    HitResult r{0.0f, 0.0f, 0.0f, -1, -1};
    int32_t ref = kEnd;
    int32_t pq[Q];
    int npend = 0;
#pragma unroll
    for (int q = 0; q < Q; q++) pq[q] = 0;
    int64_t budget = 0;
    bool more = true;
    // Ray shares: the rays are split into `shares` contiguous ranges with a counter each, on
    // lines of their own; workgroup b starts on share b % shares, and workgroups are dispatched
    // round-robin over the 8 XCDs, so with 8 shares each XCD serves one range (one hot counter
    // per XCD instead of one for the chip: 1024^2 camera rays 0.37 -> 0.20 ms; with a coherent
    // order each XCD's L2 also holds the nodes of its own region). A wave that finds its share
    // handed out goes on with the next one.
    int part = shares > 1 ? int(blockIdx.x % unsigned(shares)) : 0, tried = 0;  // wave-uniform
    for (;;) {
        const uint64_t idle = __ballot(ray < 0);
        const int nidle = __popcll(idle);
        if (more && nidle >= refill) {
            if (lane == 0) s_base[wid] = atomicAdd(counter + kTraceShareStride * part, uint32_t(nidle));
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // share `part` is [part * share_len, min(n, (part + 1) * share_len)); scalar registers
            const int64_t ps = int64_t(part) * share_len;
            const int64_t end = ps + share_len < n ? ps + share_len : n;
            const int64_t base = ps + int64_t(__builtin_amdgcn_readfirstlane(s_base[wid]));
            __builtin_amdgcn_wave_barrier();
            if (base + nidle >= end) {  // this share is handed out: go on with the next
                part = part + 1 == shares ? 0 : part + 1;
                if (++tried == shares) more = false;
            }
            if (ray < 0) {
                const int64_t q = base + __popcll(idle & below);
                if (q < end) {
                    const int64_t i = order ? int64_t(order[q]) : q;  // the ray's index in rays / hits
                    const float4 r0 = rays[2 * i], r1 = rays[2 * i + 1];
                    w = PrepareRay(r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w);
                    t = tree;
                    ref = t.node_count > 0 ? 0 : kEnd;
                    if (BATCH) {
                        const int32_t m = ray_mesh[i];
                        if (m >= 0 && m < mesh_count) {
                            const BatchTree bt = mt[m];
                            t.nodes = tree.nodes + bt.node_off * (WideNodeBytes(W) / 16);
                            t.tris = tree.tris + 3 * bt.tri_off;
                            t.node_count = bt.node_count;
                            t.prim_count = bt.prim_count;
                            ref = t.node_count > 0 ? 0 : kEnd;
                        } else {
                            ref = kEnd;  // a mesh id out of range: a miss
                        }
                    }
                    best_t = w.tmax;
                    r = HitResult{0.0f, 0.0f, 0.0f, -1, -1};
                    npend = 0;
                    st.sp = 0;
                    budget = 2 * int64_t(t.node_count) + 2;
                    ray = i;
                }
            }
        }
        if (__ballot(ray >= 0) == 0ull) {
            if (!more) break;
            continue;
        }
        if (ray >= 0) {
            for (;;) {
                if (ref < 0 && npend < Q) {
#pragma unroll
                    for (int q = Q - 1; q > 0; q--) pq[q] = pq[q - 1];
                    pq[0] = ref;
                    npend++;
                    ref = st.sp > 0 ? st.Pop() : kEnd;
                }
                const bool node = ref >= 0 && ref < t.node_count;
                if (!__any(node && npend == 0)) break;
                if (node) {
                    if (--budget < 0) {  // only a malformed tree gets here
                        ref = kEnd;
                        npend = 0;
                        st.sp = 0;
                        continue;
                    }
                    if (!WideNodeStep<W>(t, w, best_t, ref, st)) ref = st.sp > 0 ? st.Pop() : kEnd;
                }
            }
            bool done = false;
            int32_t k = 0, kend = 0;
            for (;;) {
                while (k == kend && npend > 0) {  // (an empty leaf takes the next one)
                    // the oldest pending leaf first: leaves are tested in the order the traversal
                    // reached them, whatever the other lanes held the wave to (an any-hit query then
                    // reports the first hit of its own ordered traversal)
                    int32_t oldest = pq[0];
#pragma unroll
                    for (int q = 1; q < Q; q++) oldest = npend - 1 == q ? pq[q] : oldest;
                    const int32_t packed = ~oldest;
                    npend--;
                    k = packed >> 4;
                    kend = k + (packed & 15);
                    if (kend > t.prim_count) kend = k;
                }
                const bool go = k < kend && !done;
                if (!__any(go)) break;
                if (go) {
                    done = TestTriangle<ANY>(t.tris, k, w, cull, best_t, r);
                    k++;
                }
            }
            npend = 0;
            if (ANY && done) {
                ref = kEnd;
                st.sp = 0;
            }
            if (ref >= t.node_count) {
                dpl_hit o;
                o.t = r.t;
                o.b1 = r.b1;
                o.b2 = r.b2;
                o.prim = r.prim;
                hits[ray] = o;
                ray = -1;
            }
        }
    }
}

__global__ void KResetWideCounter(uint32_t* c, unsigned long long* counters) {
#pragma unroll
    for (int k = 0; k < kTraceShares; k++) c[kTraceShareStride * k] = 0u;  // one ray counter per share
    if (counters) counters[0] = counters[1] = 0ull;
}

// A root-to-leaf path holds at most `height` inner nodes of the two-child tree, so at most
// `height` wide nodes, each pushing at most W - 1 entries.
static int32_t WideSpillStride(int width, int height) {
    const int64_t need = int64_t(width - 1) * std::max(height, 1) + 1;
    return int32_t(std::max<int64_t>(0, need - kWideStackLds));
}

static int WideGrid(int cu_count, int64_t n) {
    const int64_t need = (n + kWideBlock - 1) / kWideBlock;
    return int(std::max<int64_t>(1, std::min<int64_t>(int64_t(cu_count) * kWideWgPerCu, need)));
}

size_t TraceWideSpillBytes(int cu_count, int width, int height) {
    return size_t(cu_count) * kWideWgPerCu * kWideBlock * size_t(WideSpillStride(width, height)) * 4;
}

hipError_t LaunchTraceWide(const WideTree& tree, const dpl_ray* rays, int64_t n, uint32_t flags, dpl_hit* hits,
                           int cu_count, uint32_t* counter, int32_t* spill, unsigned long long* counters,
                           hipStream_t s, const int32_t* order) {
    if (n <= 0) return hipSuccess;
    if (n > int64_t(0x7FFFFFFF)) return hipErrorInvalidValue;  // the 32-bit ray counter never wraps
    if (tree.width != 4 && tree.width != 8) return hipErrorInvalidValue;
    hipLaunchKernelGGL(KResetWideCounter, dim3(1), dim3(1), 0, s, counter, counters);
    const float4* r4 = reinterpret_cast<const float4*>(rays);
    const int grid = WideGrid(cu_count, n);
    const int refill = TraceRefillLanes();
    const int shares = TraceSharesEnv();
    const int64_t share_len = (n + shares - 1) / shares;
#define DPL_TRACE_W(W_, COUNT_)                                                                                 \
    hipLaunchKernelGGL((KTraceWide<W_, COUNT_, false>), dim3(grid), dim3(kWideBlock), 0, s, tree, r4, n, flags,     \
                       hits, counter, spill, counters, nullptr, nullptr, 0)
#define DPL_TRACE_WR(W_, ANY_)                                                                                  \
    hipLaunchKernelGGL((KTraceWideRefill<W_, false, ANY_, kWidePending>), dim3(grid), dim3(kWideBlock), 0, s, tree, \
                       r4, n, (flags & DPL_TRACE_CULL_BACK) != 0, hits, counter, spill, nullptr, nullptr, 0, refill, \
                       order, shares, share_len)
    const bool any = flags & DPL_TRACE_ANY;
    if (counters || refill == 0) {  // the counters count the ordered traversal
        if (tree.width == 4) {
            if (counters) DPL_TRACE_W(4, true);
            else DPL_TRACE_W(4, false);
        } else {
            if (counters) DPL_TRACE_W(8, true);
            else DPL_TRACE_W(8, false);
        }
    } else if (tree.width == 4) {
        if (any) DPL_TRACE_WR(4, true);
        else DPL_TRACE_WR(4, false);
    } else {
        if (any) DPL_TRACE_WR(8, true);
        else DPL_TRACE_WR(8, false);
    }
#undef DPL_TRACE_WR
#undef DPL_TRACE_W
    return hipGetLastError();
}

hipError_t LaunchTraceWideBatched(const WideTree& tree, const BatchTree* mt, int32_t mesh_count, const dpl_ray* rays,
                                  const int32_t* ray_mesh, int64_t n, uint32_t flags, dpl_hit* hits, int cu_count,
                                  uint32_t* counter, int32_t* spill, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (n > int64_t(0x7FFFFFFF)) return hipErrorInvalidValue;  // the 32-bit ray counter never wraps
    if (tree.width != 4 && tree.width != 8) return hipErrorInvalidValue;
    hipLaunchKernelGGL(KResetWideCounter, dim3(1), dim3(1), 0, s, counter, nullptr);
    const float4* r4 = reinterpret_cast<const float4*>(rays);
    const int grid = WideGrid(cu_count, n);
    const int refill = TraceRefillLanes();
    const int shares = TraceSharesEnv();
    const int64_t share_len = (n + shares - 1) / shares;
    const bool any = flags & DPL_TRACE_ANY, cull = flags & DPL_TRACE_CULL_BACK;
    if (refill == 0) {
        if (tree.width == 4)
            hipLaunchKernelGGL((KTraceWide<4, false, true>), dim3(grid), dim3(kWideBlock), 0, s, tree, r4, n, flags,
                               hits, counter, spill, nullptr, ray_mesh, mt, mesh_count);
        else
            hipLaunchKernelGGL((KTraceWide<8, false, true>), dim3(grid), dim3(kWideBlock), 0, s, tree, r4, n, flags,
                               hits, counter, spill, nullptr, ray_mesh, mt, mesh_count);
        return hipGetLastError();
    }
#define DPL_TRACE_WB(W_, ANY_)                                                                                  \
    hipLaunchKernelGGL((KTraceWideRefill<W_, true, ANY_, kWidePending>), dim3(grid), dim3(kWideBlock), 0, s, tree,  \
                       r4, n, cull, hits, counter, spill, ray_mesh, mt, mesh_count, refill, nullptr, shares, share_len)
    if (tree.width == 4) {
        if (any) DPL_TRACE_WB(4, true);
        else DPL_TRACE_WB(4, false);
    } else {
        if (any) DPL_TRACE_WB(8, true);
        else DPL_TRACE_WB(8, false);
    }
#undef DPL_TRACE_WB
    return hipGetLastError();
}

}  // namespace dpl
