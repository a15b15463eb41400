// BVH build kernels for gfx950: scene/centroid bounds, Morton30 / EMC60 codes, PLOC
// clustering (Meister & Bittner 2017) with one fused kernel per iteration (nearest
// neighbours over an LDS window, mutual-pair merge, order-preserving compaction with a
// decoupled look-back), a single-workgroup tail that finishes the last iterations in LDS,
// a bottom-up SAH leaf collapse, and the final two-child node layout.
//
// The CPU restatement in the oracle defines the exact results; see the bit-exactness
// contract in dpl_internal.h.
#include <hip/hip_runtime.h>

#include "build_common.h"

namespace dpl {

namespace {

constexpr int kTail = 1024;        // clusters finished by the single-workgroup tail
constexpr int kTailThreads = 1024;
constexpr int kIpt = 2;            // at most this many clusters per thread in an iteration tile

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
__global__ __launch_bounds__(256) void KBounds(const float* __restrict__ tri, int64_t n, uint32_t* __restrict__ part) {
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
    // the workgroup's partial (slots: vmin 0-2, vmax 3-5, cmin 6-8, cmax 9-11, flag 12) goes
    // to its own row; KBoundsReduce combines the rows after the kernel boundary, which costs
    // far less than thousands of same-address atomics
    uint32_t* row = part + 16 * int64_t(blockIdx.x);
    if (threadIdx.x < 6) {
        int k = threadIdx.x;
        float a = s_mn[k][0], z = s_mx[k][0];
        for (int w = 1; w < int(blockDim.x / kWave); w++) {
            a = fminf(a, s_mn[k][w]);
            z = fmaxf(z, s_mx[k][w]);
        }
        int base = k < 3 ? 0 : 6, ax = k % 3;
        row[base + ax] = FloatToOrdered(a);
        row[base + 3 + ax] = FloatToOrdered(z);
    }
    if (threadIdx.x == 0) row[12] = s_bad ? 1u : 0u;
}

// One workgroup: min / max (ordered integers) and OR of the flag over the partial rows;
// each thread folds its rows, then the waves by shuffles, then the four wave results.
__device__ __forceinline__ uint32_t FoldBound(int k, uint32_t a, uint32_t b) {
    return k == 12 ? (a | b) : ((k % 6) < 3 ? min(a, b) : max(a, b));
}

__global__ __launch_bounds__(256) void KBoundsReduce(const uint32_t* __restrict__ part, int rows,
                                                     uint32_t* __restrict__ bounds) {
    uint32_t v[13];
#pragma unroll
    for (int k = 0; k < 13; k++) v[k] = (k % 6) < 3 && k < 12 ? 0xFFFFFFFFu : 0u;
    for (int r = threadIdx.x; r < rows; r += 256) {
        const uint4* row = reinterpret_cast<const uint4*>(part + 16 * int64_t(r));
        const uint4 a = row[0], b = row[1], c = row[2], d = row[3];
        const uint32_t x[13] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w, d.x};
#pragma unroll
        for (int k = 0; k < 13; k++) v[k] = FoldBound(k, v[k], x[k]);
    }
#pragma unroll
    for (int k = 0; k < 13; k++)
#pragma unroll
        for (int d = kWave / 2; d > 0; d >>= 1) v[k] = FoldBound(k, v[k], uint32_t(__shfl_xor(int(v[k]), d, kWave)));
    __shared__ uint32_t s_v[4][13];
    const int w = threadIdx.x / kWave;
    if (LaneId() == 0)
#pragma unroll
        for (int k = 0; k < 13; k++) s_v[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 13) {
        const int k = threadIdx.x;
        uint32_t r = s_v[0][k];
        for (int ww = 1; ww < 4; ww++) r = FoldBound(k, r, s_v[ww][k]);
        bounds[k] = r;
    }
}

hipError_t LaunchBounds(const BuildBuffers& b, hipStream_t s) {
    int grid = int(std::min<int64_t>(kBoundsBlocks, (b.n + 255) / 256));
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(KBounds, dim3(grid), dim3(256), 0, s, b.tri, b.n, b.bounds_part);
    hipLaunchKernelGGL(KBoundsReduce, dim3(1), dim3(256), 0, s, b.bounds_part, grid, b.bounds);
    return hipGetLastError();
}


template <int kMorton>
__global__ __launch_bounds__(256) void KCodes(const float* __restrict__ tri, int64_t n, const uint32_t* bounds,
                                               uint64_t* keys, uint32_t* vals) {
    __shared__ EmcLayoutDev sL;
    const CodeFrame f = MakeCodeFrame(Decode(bounds), kMorton);
    if (kMorton == 1) {
        if (threadIdx.x == 0) MakeEmcLayout(f.d, sL);
        __syncthreads();
        BuildEmcTables(sL, threadIdx.x, blockDim.x);
        __syncthreads();
    }
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        float v[9];
        LoadTri(tri, i, v);
        if (kMorton == 0)
            reinterpret_cast<uint32_t*>(keys)[i] = uint32_t(CodeOf<0>(v, f, sL));
        else
            keys[i] = CodeOf<1>(v, f, sL);
        vals[i] = uint32_t(i);
    }
}

hipError_t LaunchCodes(const BuildBuffers& b, int morton_type, hipStream_t s) {
    // EMC60 workgroups each build the code tables first: fewer, longer-lived workgroups
    int grid = int(std::min<int64_t>(morton_type == 1 ? 1024 : 4096, (b.n + 255) / 256));
    if (grid < 1) grid = 1;
    if (morton_type == 0)
        hipLaunchKernelGGL(KCodes<0>, dim3(grid), dim3(256), 0, s, b.tri, b.n, b.bounds, b.keys, b.vals);
    else
        hipLaunchKernelGGL(KCodes<1>, dim3(grid), dim3(256), 0, s, b.tri, b.n, b.bounds, b.keys, b.vals);
    return hipGetLastError();
}


int SobbTriples13() { return kTriples13.n; }
void SobbTriple13(int t, int* a, int* b, int* c, float* inv_det) {
    const Triple13& T = kTriples13.t[t];
    *a = T.a, *b = T.b, *c = T.c, *inv_det = T.inv_det;
}


__global__ __launch_bounds__(256) void KInitClusters(BuildBuffers b) {
    const int64_t n = b.n;
    const float pad = b.ndirs > 0 ? KdopPad(b.bounds) : 0.0f;
    for (int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < n; k += int64_t(gridDim.x) * blockDim.x)
        InitLeaf<true, true>(b, k, pad);
}

hipError_t LaunchInitClusters(const BuildBuffers& b, hipStream_t s) {
    int grid = int(std::min<int64_t>(4096, (b.n + 255) / 256));
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(KInitClusters, dim3(grid), dim3(256), 0, s, b);
    return hipGetLastError();
}


// ---------------------------------------------------------------------------------------
// one PLOC iteration over tiles of CHUNK = BLOCK * kIpt clusters; persistent over tiles
// with dynamic tile ids so the decoupled look-back always waits on running workgroups
// ---------------------------------------------------------------------------------------
constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagInc = 2ull << 62;
constexpr uint64_t kMask31 = 0x7FFFFFFFull;

// Tiles hold BLOCK * ipt clusters, ipt in {1, .., kIpt} chosen per iteration so that there are at
// least `min_tiles` tiles: small iterations spread over more workgroups, each thread then
// computing about one nearest neighbour instead of four.
// LANES > 1: BLOCK * LANES threads, the tile's nearest-neighbour searches split over LANES
// adjacent lanes each (same tiles, a quarter of the search latency for LANES = 4); the
// other phases run on threads below BLOCK, threads above own no cluster.
template <int BLOCK, int LANES>
__global__ __launch_bounds__(BLOCK * LANES) void KPlocIteration(BuildBuffers b, int iter, int R, int tail,
                                                                int max_leaf, int min_tiles) {
    constexpr int NT = BLOCK * LANES;
    // the four-lane form runs on iterations of about one tile per CU: one cluster per thread
    constexpr int IPT = LANES > 1 ? 1 : kIpt;
    constexpr int CHUNK = BLOCK * IPT;
    constexpr int WMAX = CHUNK + 4 * kMaxRadius;
    __shared__ float s_lx[WMAX], s_ly[WMAX], s_lz[WMAX], s_hx[WMAX], s_hy[WMAX], s_hz[WMAX];
    __shared__ int32_t s_id[WMAX];
    __shared__ int32_t s_nn[CHUNK + 2 * kMaxRadius];
    __shared__ uint32_t s_scan[NT / kWave + 1];
    __shared__ int s_tile;
    __shared__ uint64_t s_prefix;

    const int c = b.ccount[iter];
    // the first iteration reads the leaves from the node array: cluster k is leaf k
    const bool leaves = iter == 0;
    const PNode* __restrict__ in = leaves ? b.nodes : b.clusters[iter & 1];
    PNode* __restrict__ out = b.clusters[(iter + 1) & 1];
    const int64_t n = b.n;

    if (c <= tail) {  // done with multi-workgroup iterations: pass the clusters through
        if (blockIdx.x == 0) {
            for (int k = threadIdx.x; k < c; k += NT) {
                PNode o = in[k];
                if (leaves) o.left = k, o.right = 0;
                out[k] = o;
            }
            if (threadIdx.x == 0) {
                b.ccount[iter + 1] = c;
                b.tile_counter[(iter + 1) & 1] = 0u;
                b.results->progress = c;
            }
        }
        return;
    }
    int ipt = 1;
    while (ipt < IPT && c >= int64_t(BLOCK) * 2 * ipt * min_tiles) ipt *= 2;
    // BLOCK * ipt nearest neighbours per tile (its clusters and R on either side): every
    // thread computes exactly ipt of them
    const int chunk = BLOCK * ipt - 2 * R;
    const int T = (c + chunk - 1) / chunk;
    // the next iteration has at most Tn tiles (fewer clusters, chunks of at least
    // BLOCK - 2 R): each tile clears kc of its look-back states
    const int Tn = (c + (BLOCK - 2 * R) - 1) / (BLOCK - 2 * R);
    const int kc = (Tn + T - 1) / T;
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
        if (int(threadIdx.x) < kc && tile * kc + int(threadIdx.x) < Tn) status_next[tile * kc + threadIdx.x] = 0ull;
        if (threadIdx.x == 0 && tile == 0) b.tile_counter[(iter + 1) & 1] = 0u;
        const int a = tile * chunk;
        const int e = a + chunk < c ? a + chunk : c;
        const int w0 = a - 2 * R;
        const int wlen = (e - a) + 4 * R;
        for (int t = threadIdx.x; t < wlen; t += NT) {
            int j = w0 + t;
            if (j >= 0 && j < c) {
                const float4* p = reinterpret_cast<const float4*>(in + j);
                float4 q0 = p[0], q1 = p[1];
                s_lx[t] = q0.x, s_ly[t] = q0.y, s_lz[t] = q0.z, s_hx[t] = q0.w;
                s_hy[t] = q1.x, s_hz[t] = q1.y, s_id[t] = leaves ? j : __float_as_int(q1.z);
            }
        }
        __syncthreads();
        const int n0 = a - R < 0 ? 0 : a - R;
        const int n1 = e + R > c ? c : e + R;
        if (LANES == 1) {
            for (int i = n0 + threadIdx.x; i < n1; i += NT) s_nn[i - (a - R)] = NearestNeighbour(w, i, w0, c, R);
        } else {
            // every lane of a wave takes part in the shuffles: the loop bound is block-uniform
            const int g = int(threadIdx.x) / LANES;
            for (int i0 = n0; i0 < n1; i0 += NT / LANES) {
                const int i = i0 + g;
                const int32_t j = NearestNeighbourLanes(w, i < n1 ? i : n0, w0, c, R, LANES, i < n1);
                if (i < n1 && int(threadIdx.x) % LANES == 0) s_nn[i - (a - R)] = j;
            }
        }
        __syncthreads();

        // merge decisions for the tile's clusters (blocked: thread t owns IPT in a row)
        uint32_t sflags = 0, mflags = 0;
        int32_t nnj[IPT];
#pragma unroll
        for (int q = 0; q < IPT; q++) {
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
        const uint32_t before = BlockExclusiveScan<NT>(mine, s_scan, total);
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
        for (int q = 0; q < IPT; q++) {
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
                const float li[3] = {s_lx[ii], s_ly[ii], s_lz[ii]}, hi_[3] = {s_hx[ii], s_hy[ii], s_hz[ii]};
                const float lj[3] = {s_lx[jj], s_ly[jj], s_lz[jj]}, hj[3] = {s_hx[jj], s_hy[jj], s_hz[jj]};
                StoreEval(b, id, InnerEval(o.lo, o.hi, EvalOf(b, o.left, li, hi_), EvalOf(b, o.right, lj, hj), max_leaf));
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
        // with at least as many workgroups as tiles, the workgroups' first fetches hand out
        // every tile: leave without another round trip to the counter
        if (T <= int(gridDim.x)) break;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// The same iteration in two launches, for the large cluster counts (the host's threshold).
// In the fused kernel the workgroups of a CU move through their tiles in step (window,
// search, look-back, merge), so the search's arithmetic and the other phases' memory latency
// never overlap, and the merges' dependent loads are spread over few waves. Here no
// workgroup waits on another: KPlocNeighbours finds every cluster's nearest neighbour and
// merge decision, KPlocMerge sums each tile's offsets from the per-tile counts and merges and
// compacts, both with one thread per cluster. Same search, same merges, same order.
//
// Decision word per cluster i (in b.pairs): bits 0..29 the neighbour j, bit 30 the cluster's
// slot survives (not the higher index of a mutual pair), bit 31 it creates the merge (the
// lower index of a mutual pair). Tile counts (survivors << 32 | merges) go to the iteration's
// look-back words, which the fused kernel does not use in this iteration.
// ---------------------------------------------------------------------------------------
constexpr uint32_t kPairSurvives = 1u << 30, kPairMerges = 1u << 31, kPairIndex = kPairSurvives - 1;

// Dynamic LDS of KPlocNeighbours: R rows of `threads` pair areas, or 0 (each cluster searches
// its candidates on both sides itself). Measured on the 2M-triangle build: radius 16 with
// 512 threads (32 KiB) 2.21 -> 2.16 ms, radius 8 1.97 -> 1.94 ms, radius 4 1.80 -> 1.78 ms;
// radius 32 (64 KiB, fewer workgroups per CU) 2.80 -> 3.31 ms; radius 1 (the barrier and
// stores cost more than the areas saved) 1.56 -> 1.58 ms.
constexpr size_t kNnPairBytesMax = 32 * 1024;
constexpr int kNnPairMinRadius = 4;
__host__ __device__ constexpr size_t NnPairBytes(int R, int threads) {
    const size_t bytes = size_t(R) * size_t(threads) * sizeof(float);
    return R >= kNnPairMinRadius && bytes <= kNnPairBytesMax ? bytes : 0;
}

// Nearest neighbours of the searched clusters i = n0 + t, t < rows, with each pair's area
// computed once (the area of a union does not depend on the order of the two boxes). Cluster i
// computes SA(box_i U box_{i+k}) for k = 1..R, keeps the least in its right chain and leaves
// it in dist[(k - 1) NT + t]; after the barrier it takes its candidates i - k >= n0 from
// cluster t - k's row (the candidates before n0, which no searched cluster covers, it computes
// itself). Both chains visit their candidates in ascending j, as NnOffer's tie rule needs, and
// the two chain minima are compared with the full key: the same neighbour as NearestNeighbour.
template <int NT>
__device__ __forceinline__ void NearestNeighboursPaired(const WindowLds& w, float* __restrict__ dist, int n0, int rows,
                                                        int w0, int c, int R, int32_t* nn_out) {
    constexpr float kInf = __builtin_huge_valf();
    const int t = threadIdx.x;
    const int i = n0 + t;
    float bdl = kInf, bdr = kInf;
    int32_t bjl = -1, bjr = -1;
    if (t < rows) {
        const int ii = i - w0;
        const float alx = w.lx[ii], aly = w.ly[ii], alz = w.lz[ii], ahx = w.hx[ii], ahy = w.hy[ii], ahz = w.hz[ii];
        for (int j = i - R < 0 ? 0 : i - R; j < n0; ++j)
            NnOffer(i, j, NnArea(alx, aly, alz, ahx, ahy, ahz, w, j - w0), true, bdl, bjl);
        const int jr = i + R < c - 1 ? i + R : c - 1;
        for (int k = 1; k <= R; k += 4) {
            const int j0 = i + k, j1 = j0 + 1, j2 = j0 + 2, j3 = j0 + 3;
            // past the last cluster: re-read cluster jr (in the window) and discard the area
            const float d0 = NnArea(alx, aly, alz, ahx, ahy, ahz, w, (j0 <= jr ? j0 : jr) - w0);
            const float d1 = NnArea(alx, aly, alz, ahx, ahy, ahz, w, (j1 <= jr ? j1 : jr) - w0);
            const float d2 = NnArea(alx, aly, alz, ahx, ahy, ahz, w, (j2 <= jr ? j2 : jr) - w0);
            const float d3 = NnArea(alx, aly, alz, ahx, ahy, ahz, w, (j3 <= jr ? j3 : jr) - w0);
            float* row = dist + (k - 1) * NT + t;
            row[0] = d0;
            if (k + 1 <= R) row[NT] = d1;
            if (k + 2 <= R) row[2 * NT] = d2;
            if (k + 3 <= R) row[3 * NT] = d3;
            NnOffer(i, j0, d0, j0 <= jr, bdr, bjr);
            NnOffer(i, j1, d1, j1 <= jr, bdr, bjr);
            NnOffer(i, j2, d2, j2 <= jr, bdr, bjr);
            NnOffer(i, j3, d3, j3 <= jr, bdr, bjr);
        }
    }
    __syncthreads();
    if (t < rows) {
        // candidates i - k for k = min(R, t) down to 1; row t - k holds SA(box_{i-k} U box_i) at k
        const int kmax = R < t ? R : t;
        int k = kmax;
        for (; k >= 4; k -= 4) {
            const float d0 = dist[(k - 1) * NT + t - k], d1 = dist[(k - 2) * NT + t - k + 1];
            const float d2 = dist[(k - 3) * NT + t - k + 2], d3 = dist[(k - 4) * NT + t - k + 3];
            NnOffer(i, i - k, d0, true, bdl, bjl);
            NnOffer(i, i - k + 1, d1, true, bdl, bjl);
            NnOffer(i, i - k + 2, d2, true, bdl, bjl);
            NnOffer(i, i - k + 3, d3, true, bdl, bjl);
        }
        for (; k >= 1; --k) NnOffer(i, i - k, dist[(k - 1) * NT + t - k], true, bdl, bjl);
        if (bjl >= 0 && (bjr < 0 || NnBefore(i, bdl, bjl, bdr, bjr))) bjr = bjl;
        *nn_out = bjr;
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK * kIpt) void KPlocNeighbours(BuildBuffers b, int iter, int R, int tail) {
    constexpr int CHUNK = BLOCK * kIpt;
    constexpr int NT = CHUNK;  // one thread per searched cluster
    constexpr int WMAX = CHUNK + 4 * kMaxRadius;
    __shared__ float s_lx[WMAX], s_ly[WMAX], s_lz[WMAX], s_hx[WMAX], s_hy[WMAX], s_hz[WMAX];
    __shared__ int32_t s_nn[CHUNK + 2 * kMaxRadius];
    __shared__ uint32_t s_scan[NT / kWave + 1];

    const int c = b.ccount[iter];
    if (c <= tail) return;  // KPlocMerge passes the clusters through
    // the first iteration reads the leaves from the node array (only boxes are read here)
    const PNode* __restrict__ in = iter == 0 ? b.nodes : b.clusters[iter & 1];
    const int chunk = CHUNK - 2 * R;
    const int T = (c + chunk - 1) / chunk;
    uint64_t* counts = b.status[iter & 1];
    const WindowLds w{s_lx, s_ly, s_lz, s_hx, s_hy, s_hz};
    HIP_DYNAMIC_SHARED(float, s_dist)
    const bool paired = NnPairBytes(R, NT) != 0;  // the launch sized s_dist by the same rule
    for (int tile = blockIdx.x; tile < T; tile += gridDim.x) {
        const int a = tile * chunk;
        const int e = a + chunk < c ? a + chunk : c;
        const int w0 = a - 2 * R;
        const int wlen = (e - a) + 4 * R;
        for (int t = threadIdx.x; t < wlen; t += NT) {
            const int j = w0 + t;
            if (j >= 0 && j < c) {
                const float4* p = reinterpret_cast<const float4*>(in + j);
                const float4 q0 = p[0], q1 = p[1];
                s_lx[t] = q0.x, s_ly[t] = q0.y, s_lz[t] = q0.z, s_hx[t] = q0.w;
                s_hy[t] = q1.x, s_hz[t] = q1.y;
            }
        }
        __syncthreads();
        const int n0 = a - R < 0 ? 0 : a - R;
        const int n1 = e + R > c ? c : e + R;
        if (paired) {  // n1 - n0 <= NT: one searched cluster per thread
            NearestNeighboursPaired<NT>(w, s_dist, n0, n1 - n0, w0, c, R, s_nn + (n0 + int(threadIdx.x) - (a - R)));
        } else {
            for (int i = n0 + threadIdx.x; i < n1; i += NT) s_nn[i - (a - R)] = NearestNeighbour(w, i, w0, c, R);
        }
        __syncthreads();
        uint32_t ns = 0, nm = 0;
        const int i = a + int(threadIdx.x);
        if (i < e) {
            const int j = s_nn[i - (a - R)];
            const bool mutual = s_nn[j - (a - R)] == i;
            const bool survives = !(mutual && j < i), merges = mutual && i < j;
            b.pairs[i] = uint32_t(j) | (survives ? kPairSurvives : 0u) | (merges ? kPairMerges : 0u);
            ns = survives;
            nm = merges;
        }
        uint32_t total;
        BlockExclusiveScan<NT>(ns | (nm << 16), s_scan, total);
        if (threadIdx.x == 0) counts[tile] = (uint64_t(total & 0xFFFFu) << 32) | (total >> 16);
        __syncthreads();  // the next tile overwrites the window
    }
}

// One thread per cluster of the tile (BLOCK * kIpt threads): the merges' dependent loads
// (cluster records, collapse records, k-DOP slabs) are spread over twice the waves.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK * kIpt) __attribute__((amdgpu_waves_per_eu(6))) void KPlocMerge(BuildBuffers b, int iter, int R, int tail, int max_leaf) {
    constexpr int CHUNK = BLOCK * kIpt;
    constexpr int NT = CHUNK;
    __shared__ uint32_t s_scan[NT / kWave + 1];
    __shared__ uint64_t s_sum[NT / kWave];

    const int c = b.ccount[iter];
    // the first iteration reads the leaves from the node array: cluster k is leaf k
    const bool leaves = iter == 0;
    const PNode* __restrict__ in = leaves ? b.nodes : b.clusters[iter & 1];
    PNode* __restrict__ out = b.clusters[(iter + 1) & 1];
    const int64_t n = b.n;
    if (c <= tail) {  // done with multi-workgroup iterations: pass the clusters through
        for (int k = blockIdx.x * NT + threadIdx.x; k < c; k += gridDim.x * NT) {
            PNode o = in[k];
            if (leaves) o.left = k, o.right = 0;
            out[k] = o;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            b.ccount[iter + 1] = c;
            b.tile_counter[(iter + 1) & 1] = 0u;
            b.results->progress = c;
        }
        return;
    }
    const int chunk = CHUNK - 2 * R;
    const int T = (c + chunk - 1) / chunk;
    // the next iteration (fused or not) has at most Tn tiles; each tile clears kc of its words
    const int Tn = (c + (BLOCK - 2 * R) - 1) / (BLOCK - 2 * R);
    const int kc = (Tn + T - 1) / T;
    const uint64_t* counts = b.status[iter & 1];
    uint64_t* status_next = b.status[(iter + 1) & 1];
    const int32_t node_base = int32_t(2 * n - c);
    for (int tile = blockIdx.x; tile < T; tile += gridDim.x) {
        if (int(threadIdx.x) < kc && tile * kc + int(threadIdx.x) < Tn) status_next[tile * kc + threadIdx.x] = 0ull;
        if (threadIdx.x == 0 && tile == 0) b.tile_counter[(iter + 1) & 1] = 0u;
        // exclusive prefix of the tile: the counts of every earlier tile (survivors in the high
        // word, merges in the low word; neither sum reaches 2^32), eight loads per thread in
        // flight at once
        uint64_t part = 0;
        for (int q0 = 0; q0 < tile; q0 += 8 * NT) {
            uint64_t v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int q = q0 + u * NT + int(threadIdx.x);
                v[u] = q < tile ? counts[q] : 0ull;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) part += v[u];
        }
        part = WaveSumU64(part);
        if (LaneId() == 0) s_sum[threadIdx.x / kWave] = part;
        const int a = tile * chunk;
        const int e = a + chunk < c ? a + chunk : c;
        const int i = a + int(threadIdx.x);
        const uint32_t word = i < e ? b.pairs[i] : 0u;
        const bool survives = (word & kPairSurvives) != 0, merges = (word & kPairMerges) != 0;
        uint32_t total;
        const uint32_t before = BlockExclusiveScan<NT>(uint32_t(survives) | (uint32_t(merges) << 16), s_scan, total);
        uint64_t excl = 0;
#pragma unroll
        for (int k = 0; k < NT / kWave; k++) excl += s_sum[k];
        if (threadIdx.x == 0 && tile == T - 1) {
            const int32_t next = int32_t((excl >> 32) + (total & 0xFFFFu));
            b.ccount[iter + 1] = next;
            b.results->progress = next;
            b.results->multi_iters = iter + 1;
            b.results->flags = b.bounds[12];
        }
        if (survives) {
            PNode o = in[i];
            if (leaves) o.left = i;
            if (merges) {
                const int32_t j = int32_t(word & kPairIndex);
                PNode pj = in[j];
                if (leaves) pj.left = j;
                const NodeEval el = EvalOf(b, o.left, o.lo, o.hi), er = EvalOf(b, pj.left, pj.lo, pj.hi);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    o.lo[k] = Min2(o.lo[k], pj.lo[k]);
                    o.hi[k] = Max2(o.hi[k], pj.hi[k]);
                }
                const int32_t id = node_base + int32_t(excl & 0xFFFFFFFFull) + int32_t(before >> 16);
                o.right = pj.left;
                b.nodes[id] = o;
                b.parent[o.left] = id;
                b.parent[o.right] = id;
                StoreEval(b, id, InnerEval(o.lo, o.hi, el, er, max_leaf));
                if (b.ndirs > 3) UnionKdop(b, id, o.left, o.right);
                o.left = id;
            }
            o.right = 0;
            out[int32_t(excl >> 32) + int32_t(before & 0xFFFFu)] = o;
        }
        __syncthreads();  // s_sum and s_scan are reused by the next tile
    }
}

hipError_t LaunchPlocIterationSplit(const BuildBuffers& b, int iter, int radius, int block, int tail, int max_leaf,
                                    int grid, hipStream_t s) {
    if (grid < 1) grid = 1;
    switch (block) {
        case 128:
            hipLaunchKernelGGL(KPlocNeighbours<128>, dim3(grid), dim3(128 * kIpt), NnPairBytes(radius, 128 * kIpt), s, b, iter,
                               radius, tail);
            hipLaunchKernelGGL(KPlocMerge<128>, dim3(grid), dim3(128 * kIpt), 0, s, b, iter, radius, tail, max_leaf);
            break;
        case 512:
            hipLaunchKernelGGL(KPlocNeighbours<512>, dim3(grid), dim3(512 * kIpt), NnPairBytes(radius, 512 * kIpt), s, b, iter,
                               radius, tail);
            hipLaunchKernelGGL(KPlocMerge<512>, dim3(grid), dim3(512 * kIpt), 0, s, b, iter, radius, tail, max_leaf);
            break;
        default:
            hipLaunchKernelGGL(KPlocNeighbours<256>, dim3(grid), dim3(256 * kIpt), NnPairBytes(radius, 256 * kIpt), s, b, iter,
                               radius, tail);
            hipLaunchKernelGGL(KPlocMerge<256>, dim3(grid), dim3(256 * kIpt), 0, s, b, iter, radius, tail, max_leaf);
            break;
    }
    return hipGetLastError();
}

int PlocChunk(int block) { return block * kIpt; }
// tiles hold at least block - 2 kMaxRadius clusters
int64_t PlocStatusWords(int64_t n, int block) {
    const int64_t m = block - 2 * kMaxRadius;
    return (n + m - 1) / m + 1;
}
int PlocTail() { return kTail; }

hipError_t LaunchPlocIteration(const BuildBuffers& b, int iter, int radius, int block, int lanes, int tail,
                               int max_leaf, int grid, int min_tiles, hipStream_t s) {
    if (grid < 1) grid = 1;
    if (min_tiles < 1) min_tiles = 1;
    if (lanes == 4 && block == 256) {
        hipLaunchKernelGGL((KPlocIteration<256, 4>), dim3(grid), dim3(1024), 0, s, b, iter, radius, tail, max_leaf,
                           min_tiles);
        return hipGetLastError();
    }
    switch (block) {
        case 128:
            hipLaunchKernelGGL((KPlocIteration<128, 1>), dim3(grid), dim3(128), 0, s, b, iter, radius, tail, max_leaf,
                               min_tiles);
            break;
        case 512:
            hipLaunchKernelGGL((KPlocIteration<512, 1>), dim3(grid), dim3(512), 0, s, b, iter, radius, tail, max_leaf,
                               min_tiles);
            break;
        default:
            hipLaunchKernelGGL((KPlocIteration<256, 1>), dim3(grid), dim3(256), 0, s, b, iter, radius, tail, max_leaf,
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
    __shared__ float s_lx[kTail], s_ly[kTail], s_lz[kTail], s_hx[kTail], s_hy[kTail], s_hz[kTail];
    __shared__ int32_t s_id[kTail];
    __shared__ int32_t s_nn[kTail];
    __shared__ int32_t s_cnt[kTail], s_ml[kTail], s_ht[kTail];
    __shared__ float s_cost[kTail];
    __shared__ uint32_t s_scan[kTailThreads / kWave + 1];
    const TailLds L{s_lx, s_ly, s_lz, s_hx, s_hy, s_hz, s_id, s_nn, s_cnt, s_ml, s_ht, s_cost, s_scan};
    const int c = b.ccount[iter];
    // with no multi-workgroup iteration before it, the tail reads the leaves from the node array
    const int it = iter + TailIterations<kTailThreads, kTail>(b, iter == 0 ? b.nodes : b.clusters[iter & 1], c, R,
                                                               max_leaf, L, b.ccount + iter, &b.bounds[12], iter == 0);
    if (threadIdx.x == 0) {
        b.iter_done[0] = it;
        b.results->iter_done = it;
        b.results->tail_iters = it - iter;
    }
}

hipError_t LaunchPlocTail(const BuildBuffers& b, int iter, int radius, int max_leaf, hipStream_t s) {
    hipLaunchKernelGGL(KPlocTail, dim3(1), dim3(kTailThreads), 0, s, b, iter, radius, max_leaf);
    return hipGetLastError();
}


// The kept-node count K is the total of the keptflag scan that follows (counters[5]); the
// final tree is a full binary tree, so it has K + 1 leaves. No counting atomics here.
__global__ __launch_bounds__(256) void KClassify(BuildBuffers b, int max_leaf) {
    const int64_t n = b.n, nn = 2 * n - 1;
    int kept = 0, leaves = 0;
    for (int64_t id = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; id < nn; id += int64_t(gridDim.x) * blockDim.x)
        ClassifyNode(b, int32_t(id), max_leaf, kept, leaves);
}

// Places every triangle in its final leaf, and writes the node of every kept inner node
// (ids n .. 2n - 2; a one-leaf scene's root is id 0 when n = 1). keptflag has been scanned
// in place into exclusive ranks and leafslot into begin offsets.
__global__ __launch_bounds__(256) void KEmit(BuildBuffers b) {
    const int64_t n = b.n, nn = 2 * n - 1;
    const int32_t K = b.counters[5];
    for (int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < n; k += int64_t(gridDim.x) * blockDim.x)
        EmitTriangle<true>(b, int32_t(k));
    const int64_t first = n > 1 ? n : 0;
    for (int64_t id = first + blockIdx.x * int64_t(blockDim.x) + threadIdx.x; id < nn;
         id += int64_t(gridDim.x) * blockDim.x)
        EmitNode(b, int32_t(id), K);
}

// SOBB trees: the child slabs of every final inner node, one thread per child (lanes 2k and
// 2k + 1 fit the left and right child of node slot k, and lane 2k writes both direction
// triples): each thread waits on one child's records, not on two in turn.
// At least four waves per SIMD: the unrolled 26-DOP triple scan would otherwise be
// scheduled for instruction-level parallelism at the cost of every register.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void KFitSobb(BuildBuffers b) {
    const int32_t K = b.counters[5];
    __shared__ int4 s_tri[kMaxTriples];
    __shared__ float s_w[2 * kMaxDirs * 256];
    for (int t = threadIdx.x; t < b.ntriples; t += blockDim.x) s_tri[t] = b.triples[t];
    __syncthreads();
    const float pad = KdopPad(b.bounds);
    if (K == 0) {  // the whole scene is one leaf: the root's SOBB in both halves of node 0
        if (blockIdx.x == 0 && threadIdx.x == 0) FitNode(b, 0, 0, s_tri, pad, s_w + threadIdx.x, 256);
        return;
    }
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < 2 * int64_t(K);
         t += int64_t(gridDim.x) * blockDim.x) {
        const int64_t slot = t >> 1;
        const int side = int(t & 1);
        const int2 ch = b.kept_child[slot];
        float g[6];
        const uint32_t d = FitSobb(b, s_tri, side ? ch.y : ch.x, pad, g, s_w + threadIdx.x, 256);
        float2* o = reinterpret_cast<float2*>(b.fnodes + slot) + 3 * side;
        o[0] = make_float2(g[0], g[1]);
        o[1] = make_float2(g[2], g[3]);
        o[2] = make_float2(g[4], g[5]);
        // both lanes of the pair are in the loop together (2K is even, pairs start at even t)
        const uint32_t other = uint32_t(__shfl_xor(int(d), 1, kWave));
        if (side == 0) b.fnodes[slot].pad0 = int32_t(d | (other << 12));
    }
}

// SAH of the final tree in FP64 (Ct = Ci = 1, divided by SA(root) on the host); the node
// count is the number of kept inner nodes (0 when the whole scene is one leaf).
__global__ __launch_bounds__(256) void KSah(const FNode* nodes, const int32_t* counters, double* part,
                                            const float* dirs) {
    const int32_t count = counters[0];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        acc += SahOfNode(nodes[i], dirs);
    __shared__ double s[256];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if (int(threadIdx.x) < d) s[threadIdx.x] += s[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

hipError_t LaunchLayout(const BuildBuffers& b, int max_leaf, hipStream_t s) {
    const int64_t nn = 2 * b.n - 1;
    int grid = int(std::min<int64_t>(4096, (nn + 255) / 256));
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(KClassify, dim3(grid), dim3(256), 0, s, b, max_leaf);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // leafslot (n) and keptflag (n - 1, nothing when n = 1; its total stays 0) in one scan set
    e = LaunchExclusiveScan2(b.leafslot, b.n, b.keptflag, b.n - 1, b.scan_tmp, b.counters + 4, b.counters + 5, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(KEmit, dim3(grid), dim3(256), 0, s, b);
    return hipGetLastError();
}

// SOBB trees: fit every kept child's slab triple (after LaunchLayout); nothing for AABB trees.
hipError_t LaunchFitSobb(const BuildBuffers& b, hipStream_t s) {
    if (b.ndirs == 0) return hipSuccess;
    // at most 2 (n - 1) children: about one per thread
    const int fgrid = int(std::min<int64_t>(16384, (2 * b.n + 255) / 256));
    hipLaunchKernelGGL(KFitSobb, dim3(fgrid > 0 ? fgrid : 1), dim3(256), 0, s, b);
    return hipGetLastError();
}

// Sums the SAH partials of KSah in a fixed order (deterministic, no same-address atomics),
// then copies the build's small results into host-mapped memory.
__global__ __launch_bounds__(256) void KFinish(BuildBuffers b) {
    const double* part = reinterpret_cast<const double*>(b.bounds_part);
    double acc = 0.0;
    for (int r = threadIdx.x; r < kSahBlocks; r += 256) acc += part[r];
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) acc += __shfl_xor(acc, d, kWave);
    __shared__ double s_w[4];
    if (LaneId() == 0) s_w[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x != 0) return;
    b.sah[0] = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
    BuildResults* r = b.results;
    const int64_t root = 2 * b.n - 2;
    r->kept = b.counters[5];
    r->leaves = b.counters[5] + 1;
    r->root_height = root < b.n ? 0 : b.eval[root].height;  // a one-triangle scene: the leaf
    r->sah = b.sah[0];
    r->flags = b.bounds[12];
    const PNode nd = b.nodes[root];
    for (int a = 0; a < 3; a++) {
        r->root_lo[a] = nd.lo[a];
        r->root_hi[a] = nd.hi[a];
    }
}

hipError_t LaunchFinish(const BuildBuffers& b, hipStream_t s) {
    hipLaunchKernelGGL(KFinish, dim3(1), dim3(256), 0, s, b);
    return hipGetLastError();
}

// Partial SAH sums per workgroup into part[kSahBlocks]; KFinish adds them up.
hipError_t LaunchSah(const FNode* nodes, const int32_t* counters, double* part, const float* sobb_dirs,
                     hipStream_t s) {
    hipLaunchKernelGGL(KSah, dim3(kSahBlocks), dim3(256), 0, s, nodes, counters, part, sobb_dirs);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// exclusive scans of int32 in place: block sums, one-workgroup scan, downsweep
// ---------------------------------------------------------------------------------------
namespace {
constexpr int kScanBlock = 256;
constexpr int kScanIpt = 8;
constexpr int kScanTile = kScanBlock * kScanIpt;
}  // namespace

size_t ScanTempInts(int64_t n) { return size_t((n + kScanTile - 1) / kScanTile) + 1; }

// Sum of tile `tile` of in[0, n), returned to thread 0.
__device__ __forceinline__ int32_t ScanTileSum(const int32_t* in, int64_t n, int64_t tile, int32_t* s) {
    const int64_t base = tile * kScanTile;
    int32_t acc = 0;
    for (int k = threadIdx.x; k < kScanTile; k += kScanBlock) {
        int64_t i = base + k;
        if (i < n) acc += in[i];
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int d = kScanBlock / 2; d > 0; d >>= 1) {
        if (int(threadIdx.x) < d) s[threadIdx.x] += s[threadIdx.x + d];
        __syncthreads();
    }
    const int32_t r = s[0];
    __syncthreads();
    return r;
}

// One exclusive scan of the tile sums sums[0, m) in place, its total to *total.
__device__ __forceinline__ void ScanSums(int32_t* sums, int64_t m, int32_t* total) {
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

// Tile `tile` of the exclusive scan of in[0, n) into out, from the tile's base `sum`.
__device__ __forceinline__ void ScanTileDown(const int32_t* in, int32_t* out, int64_t n, int64_t tile, int32_t sum,
                                             uint32_t* s_scan) {
    const int64_t base = tile * kScanTile + int64_t(threadIdx.x) * kScanIpt;
    int32_t v[kScanIpt];
    int32_t acc = 0;
#pragma unroll
    for (int q = 0; q < kScanIpt; q++) {
        int64_t i = base + q;
        v[q] = i < n ? in[i] : 0;
        acc += v[q];
    }
    uint32_t tot;
    int32_t run = sum + int32_t(BlockExclusiveScan<kScanBlock>(uint32_t(acc), s_scan, tot));
#pragma unroll
    for (int q = 0; q < kScanIpt; q++) {
        int64_t i = base + q;
        if (i < n) out[i] = run;
        run += v[q];
    }
}

// Two in-place exclusive scans (x[0, nx) and y[0, ny), ny <= nx) in the same three launches:
// the tile sums of x in tmp[0, blocks), of y in tmp[blocks, 2 blocks); workgroup 0 of the
// middle launch scans x's sums, workgroup 1 y's.
__global__ __launch_bounds__(kScanBlock) void KScanReduce2(const int32_t* x, int64_t nx, const int32_t* y, int64_t ny,
                                                           int32_t* sums, int64_t blocks) {
    __shared__ int32_t s[kScanBlock];
    const int32_t tx = ScanTileSum(x, nx, blockIdx.x, s);
    const int32_t ty = ScanTileSum(y, ny, blockIdx.x, s);
    if (threadIdx.x == 0) {
        sums[blockIdx.x] = tx;
        sums[blocks + blockIdx.x] = ty;
    }
}

__global__ __launch_bounds__(1024) void KScanSums2(int32_t* sums, int64_t blocks, int32_t* total_x, int32_t* total_y) {
    if (blockIdx.x == 0)
        ScanSums(sums, blocks, total_x);
    else
        ScanSums(sums + blocks, blocks, total_y);
}

__global__ __launch_bounds__(kScanBlock) void KScanDown2(int32_t* x, int64_t nx, int32_t* y, int64_t ny,
                                                         const int32_t* sums, int64_t blocks) {
    __shared__ uint32_t s_scan[kScanBlock / kWave + 1];
    ScanTileDown(x, x, nx, blockIdx.x, sums[blockIdx.x], s_scan);
    ScanTileDown(y, y, ny, blockIdx.x, sums[blocks + blockIdx.x], s_scan);
}

hipError_t LaunchExclusiveScan2(int32_t* x, int64_t nx, int32_t* y, int64_t ny, int32_t* tmp, int32_t* total_x,
                                int32_t* total_y, hipStream_t s) {
    const int64_t blocks = std::max<int64_t>(1, (nx + kScanTile - 1) / kScanTile);
    hipLaunchKernelGGL(KScanReduce2, dim3(unsigned(blocks)), dim3(kScanBlock), 0, s, x, nx, y, ny, tmp, blocks);
    hipLaunchKernelGGL(KScanSums2, dim3(2), dim3(1024), 0, s, tmp, blocks, total_x, total_y);
    hipLaunchKernelGGL(KScanDown2, dim3(unsigned(blocks)), dim3(kScanBlock), 0, s, x, nx, y, ny, tmp, blocks);
    return hipGetLastError();
}

}  // namespace dpl
