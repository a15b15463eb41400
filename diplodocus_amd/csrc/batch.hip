// Batched BVH build: many independent meshes in ONE persistent launch. Every workgroup takes
// whole meshes from a work queue (largest first, so the longest jobs start early) and runs
// all stages of the single-scene build for its mesh: bounds, Morton30 / EMC60 codes, a
// workgroup LSD radix sort, leaves, PLOC (chunked iterations over global memory while the
// mesh has more clusters than fit in LDS, then every remaining iteration LDS-resident), the
// leaf collapse and the final layout. Each mesh's result equals the single-scene build of
// that mesh alone: the same codes, stable order, merge list (node ids local to the mesh),
// final nodes and triangle order. AABB node bounds only.
#include <hip/hip_runtime.h>

#include "build_common.h"

namespace dpl {

namespace {

constexpr int kBTail = 1024;  // clusters resident in LDS
constexpr int kBChunk = 1024; // clusters per chunk of a chunked iteration
constexpr int kBWin = kBChunk + 4 * kMaxRadius;   // window [a - 2R, e + 2R)
constexpr int kBNn = kBChunk + 2 * kMaxRadius;    // nearest neighbours of [a - R, e + R)
constexpr int kMaxRadix = 256;

// LDS of one workgroup of kBT threads, reused by the stages (one stage at a time).
template <int kBT>
struct alignas(16) BatchLds {
    static constexpr int kBWaves = kBT / kWave;
    union {
        struct {
            float lx[kBWin], ly[kBWin], lz[kBWin], hx[kBWin], hy[kBWin], hz[kBWin];
            int32_t id[kBWin];
            int32_t nn[kBNn];
        } win;
        struct {
            float lx[kBTail], ly[kBTail], lz[kBTail], hx[kBTail], hy[kBTail], hz[kBTail];
            int32_t id[kBTail], nn[kBTail], cnt[kBTail], ml[kBTail], ht[kBTail];
            float cost[kBTail];
        } tail;
        struct {
            uint32_t hist[kMaxRadix];   // digit histogram of the pass
            uint32_t base[kMaxRadix];   // running scatter base per digit
            uint32_t total[kMaxRadix];  // digit counts of the current chunk
            uint32_t wcnt[kBWaves * kMaxRadix];
        } sort;
        double sah[kBT];
    } u;
    uint32_t scan[kBWaves + 1];
    float red_mn[6][kBWaves], red_mx[6][kBWaves];
    DecodedBounds bounds;
    EmcLayoutDev emc;
    int job;
    int bad;
};

__device__ __forceinline__ BuildBuffers MeshView(const BatchArgs& A, int m, int64_t off, int64_t n) {
    BuildBuffers v{};
    const int64_t nb = 2 * off - m;  // first PLOC node of the mesh
    v.n = n;
    v.tri = A.tri + 9 * off;
    v.keys = A.keys + off;
    v.keys_alt = A.keys_alt + off;
    v.vals = A.vals + off;
    v.vals_alt = A.vals_alt + off;
    v.nodes = A.nodes + nb;
    v.parent = A.parent + nb;
    v.eval = A.eval + nb;
    v.state = A.state + nb;
    v.clusters[0] = A.clusters0 + off;
    v.clusters[1] = A.clusters1 + off;
    v.leafslot = A.leafslot + off;
    v.keptflag = A.keptflag + off;
    v.fnodes = A.fnodes + A.fnode_off[m];
    v.tri_ordered = A.tri_ordered + 3 * off;
    v.prim_order = A.prim_order + off;
    v.ndirs = 0;
    return v;
}

// Vertex and centroid bounds of the mesh (fminf/fmaxf of canonical coordinates: exact and
// equal to the single-scene reduction); flags non-finite coordinates.
template <int kBT>
__device__ void MeshBounds(const BuildBuffers& v, BatchLds<kBT>& S) {
    constexpr int kBWaves = kBT / kWave;
    float mn[6], mx[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        mn[k] = __int_as_float(0x7f800000);
        mx[k] = -__int_as_float(0x7f800000);
    }
    bool bad = false;
    for (int64_t i = threadIdx.x; i < v.n; i += kBT) {
        float t[9], lo[3], hi[3], c[3];
        LoadTri(v.tri, i, t);
#pragma unroll
        for (int k = 0; k < 9; k++) bad |= !isfinite(t[k]);
        TriBox(t, lo, hi);
        Centroid(t, c);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            mn[a] = fminf(mn[a], lo[a]);
            mx[a] = fmaxf(mx[a], hi[a]);
            mn[3 + a] = fminf(mn[3 + a], c[a]);
            mx[3 + a] = fmaxf(mx[3 + a], c[a]);
        }
    }
    if (threadIdx.x == 0) S.bad = 0;
    const int wid = threadIdx.x / kWave, lane = LaneId();
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const float a = WaveMin(mn[k]), z = WaveMax(mx[k]);
        if (lane == 0) {
            S.red_mn[k][wid] = a;
            S.red_mx[k][wid] = z;
        }
    }
    __syncthreads();
    if (bad) atomicOr(&S.bad, 1);
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        float a = S.red_mn[k][0], z = S.red_mx[k][0];
        for (int w = 1; w < kBWaves; w++) {
            a = fminf(a, S.red_mn[k][w]);
            z = fmaxf(z, S.red_mx[k][w]);
        }
        // round trip through the ordered-int encoding of the single-scene reduction
        a = OrderedToFloat(FloatToOrdered(a));
        z = OrderedToFloat(FloatToOrdered(z));
        if (k < 3) {
            S.bounds.vmin[k] = a;
            S.bounds.vmax[k] = z;
        } else {
            S.bounds.cmin[k - 3] = a;
            S.bounds.cmax[k - 3] = z;
        }
    }
    __syncthreads();
}

// Stable LSD radix sort of (key, value) pairs by the low `bits` bits, digit width db (4 or
// 8), through global memory: per pass one LDS histogram, then chunks of kBT keys scattered in
// index order (rank within a wave from digit ballots, wave offsets from LDS counts). Returns
// true when the sorted pairs ended in the alternate buffers.
template <int kBT, typename K>
__device__ bool WgRadixSort(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt, int64_t n, int bits, int db,
                            BatchLds<kBT>& S) {
    constexpr int kBWaves = kBT / kWave;
    const int passes = (bits + db - 1) / db;
    const uint32_t radix = 1u << db, mask = radix - 1u;
    const int wid = threadIdx.x / kWave, lane = LaneId();
    K* src = keys;
    uint32_t* srcv = vals;
    K* dst = keys_alt;
    uint32_t* dstv = vals_alt;
    bool alt = false;
    auto& so = S.u.sort;
    for (int p = 0; p < passes; p++) {
        const int shift = p * db;
        for (uint32_t d = threadIdx.x; d < radix; d += kBT) so.hist[d] = 0u;
        __syncthreads();
        for (int64_t i = threadIdx.x; i < n; i += kBT) atomicAdd(&so.hist[uint32_t(src[i] >> shift) & mask], 1u);
        __syncthreads();
        {
            uint32_t total;
            const uint32_t h = threadIdx.x < radix ? so.hist[threadIdx.x] : 0u;
            const uint32_t before = BlockExclusiveScan<kBT>(h, S.scan, total);
            if (threadIdx.x < radix) so.base[threadIdx.x] = before;
        }
        __syncthreads();
        for (int64_t a = 0; a < n; a += kBT) {
            for (int k = threadIdx.x; k < kBWaves * int(radix); k += kBT) so.wcnt[k] = 0u;
            __syncthreads();
            const int64_t i = a + threadIdx.x;
            const bool valid = i < n;
            const K key = valid ? src[i] : K(0);
            const uint32_t val = valid ? srcv[i] : 0u;
            const uint32_t d = uint32_t(key >> shift) & mask;
            uint64_t same = __ballot(valid);
            for (int bit = 0; bit < db; bit++) {
                const uint64_t ones = __ballot(((d >> bit) & 1u) != 0u);
                same &= ((d >> bit) & 1u) ? ones : ~ones;
            }
            const uint32_t rank = uint32_t(__popcll(same & ((1ull << lane) - 1ull)));
            if (valid && rank == 0) so.wcnt[wid * radix + d] = uint32_t(__popcll(same));
            __syncthreads();
            if (threadIdx.x < radix) {
                uint32_t acc = 0;
                for (int w = 0; w < kBWaves; w++) {
                    const uint32_t t = so.wcnt[w * radix + threadIdx.x];
                    so.wcnt[w * radix + threadIdx.x] = acc;
                    acc += t;
                }
                so.total[threadIdx.x] = acc;
            }
            __syncthreads();
            if (valid) {
                const uint32_t pos = so.base[d] + so.wcnt[wid * radix + d] + rank;
                dst[pos] = key;
                dstv[pos] = val;
            }
            __syncthreads();
            if (threadIdx.x < radix) so.base[threadIdx.x] += so.total[threadIdx.x];
            __syncthreads();
        }
        K* tk = src;
        src = dst;
        dst = tk;
        uint32_t* tv = srcv;
        srcv = dstv;
        dstv = tv;
        alt = !alt;
    }
    return alt;
}

// One PLOC iteration over c > kBTail clusters in chunks of kBChunk, in order, so that the
// compaction offsets are running sums (no look-back). Returns the clusters left.
template <int kBT>
__device__ int ChunkedIteration(const BuildBuffers& v, const PNode* in, PNode* out, int c, int R, int max_leaf,
                                BatchLds<kBT>& S) {
    constexpr int kBIpt = kBChunk / kBT;
    auto& W = S.u.win;
    const WindowLds w{W.lx, W.ly, W.lz, W.hx, W.hy, W.hz};
    const int32_t node_base = int32_t(2 * v.n - c);
    int32_t s_run = 0, m_run = 0;
    for (int a = 0; a < c; a += kBChunk) {
        const int e = a + kBChunk < c ? a + kBChunk : c;
        const int w0 = a - 2 * R;
        const int wlen = (e - a) + 4 * R;
        for (int t = threadIdx.x; t < wlen; t += kBT) {
            const int j = w0 + t;
            if (j >= 0 && j < c) {
                const float4* p = reinterpret_cast<const float4*>(in + j);
                const float4 q0 = p[0], q1 = p[1];
                W.lx[t] = q0.x, W.ly[t] = q0.y, W.lz[t] = q0.z, W.hx[t] = q0.w;
                W.hy[t] = q1.x, W.hz[t] = q1.y, W.id[t] = __float_as_int(q1.z);
            }
        }
        __syncthreads();
        const int n0 = a - R < 0 ? 0 : a - R;
        const int n1 = e + R > c ? c : e + R;
        for (int i = n0 + threadIdx.x; i < n1; i += kBT) W.nn[i - (a - R)] = NearestNeighbour(w, i, w0, c, R);
        __syncthreads();
        uint32_t sflags = 0, mflags = 0;
        int32_t nnj[kBIpt];
#pragma unroll
        for (int q = 0; q < kBIpt; q++) {
            const int i = a + threadIdx.x * kBIpt + q;
            nnj[q] = -1;
            if (i < e) {
                const int j = W.nn[i - (a - R)];
                const bool mutual = W.nn[j - (a - R)] == i;
                nnj[q] = j;
                if (!(mutual && j < i)) sflags |= 1u << q;
                if (mutual && i < j) mflags |= 1u << q;
            }
        }
        const uint32_t mine = uint32_t(__popc(sflags)) | (uint32_t(__popc(mflags)) << 16);
        uint32_t total;
        const uint32_t before = BlockExclusiveScan<kBT>(mine, S.scan, total);
        int32_t spos = s_run + int32_t(before & 0xFFFFu);
        int32_t mpos = node_base + m_run + int32_t(before >> 16);
#pragma unroll
        for (int q = 0; q < kBIpt; q++) {
            const int i = a + threadIdx.x * kBIpt + q;
            if (!(sflags & (1u << q))) continue;
            const int ii = i - w0;
            PNode o;
            if (mflags & (1u << q)) {
                const int jj = nnj[q] - w0;
                o.lo[0] = Min2(W.lx[ii], W.lx[jj]);
                o.lo[1] = Min2(W.ly[ii], W.ly[jj]);
                o.lo[2] = Min2(W.lz[ii], W.lz[jj]);
                o.hi[0] = Max2(W.hx[ii], W.hx[jj]);
                o.hi[1] = Max2(W.hy[ii], W.hy[jj]);
                o.hi[2] = Max2(W.hz[ii], W.hz[jj]);
                const int32_t id = mpos++;
                o.left = W.id[ii];
                o.right = W.id[jj];
                v.nodes[id] = o;
                v.parent[o.left] = id;
                v.parent[o.right] = id;
                StoreEval(v, id, InnerEval(o.lo, o.hi, LoadEval(v, o.left), LoadEval(v, o.right), max_leaf));
                o.left = id;
                o.right = 0;
            } else {
                o.lo[0] = W.lx[ii], o.lo[1] = W.ly[ii], o.lo[2] = W.lz[ii];
                o.hi[0] = W.hx[ii], o.hi[1] = W.hy[ii], o.hi[2] = W.hz[ii];
                o.left = W.id[ii];
                o.right = 0;
            }
            out[spos++] = o;
        }
        s_run += int32_t(total & 0xFFFFu);
        m_run += int32_t(total >> 16);
        __syncthreads();
    }
    return s_run;
}

// In-place exclusive scan of arr[0..len) by the workgroup; returns the total.
template <int kBT>
__device__ int32_t WgExclusiveScan(int32_t* arr, int64_t len, BatchLds<kBT>& S) {
    int32_t run = 0;
    for (int64_t a = 0; a < len; a += kBT) {
        const int64_t i = a + threadIdx.x;
        const uint32_t x = i < len ? uint32_t(arr[i]) : 0u;
        uint32_t total;
        const uint32_t before = BlockExclusiveScan<kBT>(x, S.scan, total);
        if (i < len) arr[i] = run + int32_t(before);
        run += int32_t(total);
        __syncthreads();
    }
    return run;
}

// At least four waves per SIMD: with 512-thread workgroups that keeps two resident per CU
// (128 VGPRs, a few spilled) instead of one at 160 VGPRs, 25 % faster on 4096 meshes. The
// 256-thread kernel is held to three workgroups per CU by its 52 KB of LDS either way.
template <int kBT>
__global__ __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(kBT == 512 ? 4 : 1))) void KBatchBuild(BatchArgs A) {
    __shared__ BatchLds<kBT> S;
    for (;;) {
        if (threadIdx.x == 0) S.job = int(atomicAdd(A.queue, 1u));
        __syncthreads();
        const int job = __builtin_amdgcn_readfirstlane(S.job);  // uniform: keeps the mesh view in SGPRs
        __syncthreads();
        if (job >= A.mesh_count) break;
        const int m = A.order[job];
        const int64_t off = A.tri_off[m], n = A.tri_off[m + 1] - off;
        BuildBuffers v = MeshView(A, m, off, n);
        const int64_t nn = 2 * n - 1;

        // ---- bounds and codes ----
        MeshBounds(v, S);
        const bool bad = S.bad != 0;
        if (A.morton_type == 1 && threadIdx.x == 0) MakeEmcLayout(S.bounds, S.emc);
        __syncthreads();
        if (A.morton_type == 1) {
            BuildEmcTables(S.emc, threadIdx.x, kBT);
            __syncthreads();
        }
        const CodeFrame f = MakeCodeFrame(S.bounds, A.morton_type);
        for (int64_t i = threadIdx.x; i < n; i += kBT) {
            float t[9];
            LoadTri(v.tri, i, t);
            if (A.morton_type == 0)
                reinterpret_cast<uint32_t*>(v.keys)[i] = uint32_t(CodeOf<0>(t, f, S.emc));
            else
                v.keys[i] = CodeOf<1>(t, f, S.emc);
            v.vals[i] = uint32_t(i);
        }
        __syncthreads();

        // ---- sort ----
        bool alt;
        if (A.morton_type == 0)
            alt = WgRadixSort<kBT, uint32_t>(reinterpret_cast<uint32_t*>(v.keys), v.vals,
                                        reinterpret_cast<uint32_t*>(v.keys_alt), v.vals_alt, n, 30, A.digit_bits, S);
        else
            alt = WgRadixSort<kBT, uint64_t>(v.keys, v.vals, v.keys_alt, v.vals_alt, n, 60, A.digit_bits, S);
        if (alt) v.vals = v.vals_alt;

        // ---- leaves ----
        for (int64_t k = threadIdx.x; k < n; k += kBT) {
            InitLeaf(v, k, 0.0f);
            v.leafslot[k] = 0;
        }
        if (threadIdx.x == 0) v.parent[nn - 1] = -1;
        __syncthreads();

        // ---- PLOC ----
        uint32_t* flags = &A.results[m].flags;
        int c = int(n), parity = 0, iters = 0;
        bool failed = false;
        while (c > kBTail) {
            const int cn = ChunkedIteration(v, v.clusters[parity], v.clusters[parity ^ 1], c, A.radius, A.max_leaf, S);
            if (cn >= c) {  // cannot happen for a symmetric distance; never loop on it
                if (threadIdx.x == 0) atomicOr(flags, 2u);
                failed = true;
                break;
            }
            c = cn;
            parity ^= 1;
            iters++;
        }
        if (!failed && c > 1) {
            auto& T = S.u.tail;
            const TailLds L{T.lx, T.ly, T.lz, T.hx, T.hy, T.hz, T.id, T.nn, T.cnt, T.ml, T.ht, T.cost, S.scan};
            iters += TailIterations<kBT, kBTail>(v, v.clusters[parity], c, A.radius, A.max_leaf, L, nullptr, flags);
        }
        __syncthreads();
        if (failed || (__hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & 2u)) {
            __syncthreads();
            continue;  // the host reports the mesh's flag; its layout is not built
        }

        // ---- collapse and layout ----
        int kept = 0, leaves = 0;
        for (int64_t id = threadIdx.x; id < nn; id += kBT) ClassifyNode(v, int32_t(id), A.max_leaf, kept, leaves);
        __syncthreads();
        const int32_t K = n > 1 ? WgExclusiveScan(v.keptflag, n - 1, S) : 0;
        WgExclusiveScan(v.leafslot, n, S);
        uint32_t lt;
        BlockExclusiveScan<kBT>(uint32_t(leaves), S.scan, lt);
        __syncthreads();
        for (int64_t k = threadIdx.x; k < n; k += kBT) EmitTriangle(v, int32_t(k));
        for (int64_t id = (n > 1 ? n : 0) + threadIdx.x; id < nn; id += kBT) EmitNode(v, int32_t(id), K);
        __syncthreads();

        // ---- SAH terms and the mesh's results ----
        double acc = 0.0;
        for (int i = threadIdx.x; i < K; i += kBT) acc += SahOfNode(v.fnodes[i], nullptr);
        S.u.sah[threadIdx.x] = acc;
        __syncthreads();
        for (int d = kBT / 2; d > 0; d >>= 1) {
            if (int(threadIdx.x) < d) S.u.sah[threadIdx.x] += S.u.sah[threadIdx.x + d];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            MeshResult& r = A.results[m];
            r.kept = K;
            r.leaves = int32_t(lt);
            r.root_height = v.eval[nn - 1].height;
            r.iterations = iters;
            r.sah = S.u.sah[0];
            if (bad) r.flags |= 1u;
            const PNode root = v.nodes[nn - 1];
            for (int a = 0; a < 3; a++) {
                r.root_lo[a] = root.lo[a];
                r.root_hi[a] = root.hi[a];
            }
        }
        __syncthreads();
    }
}

}  // namespace

// Resident workgroups per CU: 256 threads: registers (182 VGPRs) allow 2; 512 threads: 2 at
// four waves per SIMD.
int BatchWgPerCu(int block) { return block == 256 ? 2 : 2; }

hipError_t LaunchBatchBuild(const BatchArgs& A, int block, int grid, hipStream_t s) {
    if (grid < 1) grid = 1;
    if (block == 256)
        hipLaunchKernelGGL(KBatchBuild<256>, dim3(grid), dim3(256), 0, s, A);
    else
        hipLaunchKernelGGL(KBatchBuild<512>, dim3(grid), dim3(512), 0, s, A);
    return hipGetLastError();
}

}  // namespace dpl
