// Internal declarations of libdpl: device record layouts, device helpers shared by the
// build, sort and trace kernels, and the host-side launchers each .hip file exports.
//
// Bit-exactness contract: every stage whose output is compared bit for bit against the
// CPU restatement (bounds, codes, order, PLOC merges, collapse, layout) uses only IEEE
// add/sub/mul/div/sqrt in a fixed order, min/max in the std::min/std::max select form,
// and is compiled with -ffp-contract=off (no FMA contraction) and IEEE denormals.
#pragma once
#include <cstdlib>

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/dpl/dpl.h"

namespace dpl {

constexpr int kWave = 64;
constexpr int kMaxRadius = 32;
constexpr int kMaxLeaf = 15;

// PLOC node / cluster record (32 B). For a leaf: left = input triangle, right = -1.
// For a cluster: left = node id.
struct alignas(16) PNode {
    float lo[3], hi[3];
    int32_t left, right;
};
static_assert(sizeof(PNode) == 32, "PNode is 32 bytes");

// Final traversal node (64 B): both child boxes plus child references. A reference r >= 0
// is an inner node; r < 0 is a leaf with ~r = (prim_begin << 4) | prim_count.
struct alignas(16) FNode {
    float lx0, lx1, ly0, ly1;
    float rx0, rx1, ry0, ry1;
    float lz0, lz1, rz0, rz1;
    int32_t left, right, pad0, pad1;
};
static_assert(sizeof(FNode) == 64, "FNode is 64 bytes");

// ---------------------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ float Min2(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float Max2(float a, float b) { return (a < b) ? b : a; }
// -0.0 -> +0.0; every loaded vertex coordinate goes through it.
__device__ __forceinline__ float Canon(float x) { return x + 0.0f; }

__device__ __forceinline__ uint32_t FloatToOrdered(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float OrderedToFloat(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// The reference's avalanche hash (src/util/util.h:24-32).
__host__ __device__ __forceinline__ uint32_t HashU32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// Priority of an unordered cluster pair, used to break exact distance ties in the PLOC
// nearest-neighbour search (identical for (i,j) and (j,i)).
__device__ __forceinline__ uint32_t PairHash(int32_t i, int32_t j) {
    uint32_t lo = uint32_t(i < j ? i : j), hi = uint32_t(i < j ? j : i);
    return HashU32(lo * 0x9E3779B1u ^ hi);
}

// A size that overflows is clamped to the largest float: the area is never NaN, so the
// nearest-neighbour keys (area, pair hash, index) are totally ordered.
__device__ __forceinline__ float SurfaceAreaUnion(float alx, float aly, float alz, float ahx, float ahy, float ahz,
                                                  float blx, float bly, float blz, float bhx, float bhy, float bhz) {
    constexpr float kMax = 3.40282347e38f;
    float sx = Min2(Max2(ahx, bhx) - Min2(alx, blx), kMax);
    float sy = Min2(Max2(ahy, bhy) - Min2(aly, bly), kMax);
    float sz = Min2(Max2(ahz, bhz) - Min2(alz, blz), kMax);
    return 2.0f * (sx * sy + sy * sz + sx * sz);
}

__device__ __forceinline__ float SurfaceAreaBox(const float* lo, const float* hi) {
    float sx = hi[0] - lo[0], sy = hi[1] - lo[1], sz = hi[2] - lo[2];
    return 2.0f * (sx * sy + sy * sz + sx * sz);
}

__device__ __forceinline__ int LaneId() { return __lane_id(); }

// Inclusive scan of one uint32 per lane across the 64-lane wave.
__device__ __forceinline__ uint32_t WaveInclusiveScan(uint32_t v) {
    const int lane = LaneId();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        uint32_t o = __shfl_up(v, d, kWave);
        if (lane >= d) v += o;
    }
    return v;
}

__device__ __forceinline__ uint64_t WaveSumU64(uint64_t v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        const uint32_t lo = __shfl_xor(uint32_t(v), d, kWave);
        const uint32_t hi = __shfl_xor(uint32_t(v >> 32), d, kWave);
        v += (uint64_t(hi) << 32) | lo;
    }
    return v;
}

// Decoupled look-back over 64-bit tile states (2 flag bits on top: 1 = aggregate, 2 =
// inclusive prefix; values add in the low 62 bits), with one wave reading 64 predecessors
// per round: lane l reads tile - 1 - l, the nearest inclusive state ends the walk, and a
// round in which any tile up to that one has not published yet is retried. Called by the
// whole first wave of the workgroup; every lane returns the exclusive prefix of `tile`.
// Tile 0 publishes its inclusive state before it could wait, so the walk always ends.
__device__ __forceinline__ uint64_t WaveLookBack(uint64_t* status, int tile) {
    const int lane = int(threadIdx.x) & (kWave - 1);
    constexpr uint64_t kFlags = 3ull << 62, kInc = 2ull << 62;
    uint64_t excl = 0;
    int top = tile - 1;
    for (;;) {
        const int q = top - lane;
        const uint64_t v = q >= 0 ? __hip_atomic_load(&status[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : kInc;
        const uint64_t flag = v & kFlags;
        const uint64_t inc_mask = __ballot(flag == kInc);
        const uint64_t zero_mask = __ballot(flag == 0);
        const int first_inc = inc_mask ? __ffsll(static_cast<unsigned long long>(inc_mask)) - 1 : kWave;
        const uint64_t need = first_inc >= kWave - 1 ? ~0ull : ((2ull << first_inc) - 1);  // lanes 0..first_inc
        if (zero_mask & need) {
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        excl += WaveSumU64(lane <= first_inc ? (v & ~kFlags) : 0ull);
        if (first_inc < kWave) break;
        top -= kWave;
    }
    return excl;
}

__device__ __forceinline__ float WaveMin(float v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v = fminf(v, __shfl_xor(v, d, kWave));
    return v;
}
__device__ __forceinline__ float WaveMax(float v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, kWave));
    return v;
}

// Exclusive scan across a block of BLOCK threads; `total` receives the block sum.
// `scratch` needs BLOCK/64 + 1 uint32 of LDS. Contains __syncthreads().
template <int BLOCK>
__device__ __forceinline__ uint32_t BlockExclusiveScan(uint32_t v, uint32_t* scratch, uint32_t& total) {
    constexpr int W = BLOCK / kWave;
    const int lane = LaneId();
    const int wid = threadIdx.x / kWave;
    uint32_t inc = WaveInclusiveScan(v);
    if (lane == kWave - 1) scratch[wid] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < W; w++) {
            uint32_t t = scratch[w];
            scratch[w] = acc;
            acc += t;
        }
        scratch[W] = acc;
    }
    __syncthreads();
    total = scratch[W];
    uint32_t r = scratch[wid] + inc - v;
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------------------
// launchers (host side); every one enqueues on `s` and returns the launch error
// ---------------------------------------------------------------------------------------
// Small build results the kernels write straight into host-mapped memory (no copies).
struct BuildResults {
    int32_t progress;     // clusters after the last finished multi-workgroup iteration
    uint32_t flags;       // 1: non-finite input, 2: an iteration without a merge
    int32_t multi_iters;  // multi-workgroup iterations that merged (not pass-through)
    int32_t tail_iters;   // iterations of the single-workgroup tail
    int32_t iter_done;    // index one past the last recorded cluster count
    int32_t kept;         // final inner nodes (0: the whole scene is one leaf)
    int32_t leaves;       // final leaves
    int32_t root_height;  // longest root-to-leaf path of the PLOC tree
    double sah;           // FP64 sum of the children's SAH terms over the final nodes
    float root_lo[3], root_hi[3];
};

// Fused collapse record of a PLOC node, one 16-byte load or store: triangle count with the
// collapsed flag in the sign bit, SAH cost, smallest sorted position below it, height.
struct alignas(16) EvalRec {
    int32_t count_col;
    float cost;
    int32_t minleaf;
    int32_t height;
};

struct BuildBuffers {
    int64_t n;
    const float* tri;      // n x 9 (device)
    uint32_t* bounds;      // 13: ordered vmin3 vmax3 cmin3 cmax3, error flag
    uint32_t* bounds_part; // kBoundsBlocks x 16: per-workgroup partial bounds (same order)
    uint64_t* keys;        // n (u32 view for Morton30)
    uint32_t* vals;        // n
    uint64_t* keys_alt;    // n
    uint32_t* vals_alt;    // n
    uint32_t* pairs;       // n: per-cluster merge decisions of the two-launch PLOC iteration
                           // (the sort's spare value buffer once the sort is done)
    PNode* nodes;          // 2n-1
    int32_t* parent;       // 2n-1
    PNode* clusters[2];    // n each
    int32_t* ccount;       // n+2: clusters at the start of each iteration
    uint64_t* status[2];   // look-back words per tile
    uint32_t* tile_counter;  // 2
    int32_t* iter_done;    // 1: total iterations
    EvalRec* eval;         // 2n-1: fused collapse record
    uint8_t* state;        // 2n-1: final-layout class (0 inside a leaf, 1 final leaf, 2 kept inner)
    BuildResults* results;   // host-mapped
    int32_t* leafslot;     // n: count of the final leaf whose minimum sorted position is k
    int32_t* keptflag;     // n: internal node kept as a final inner node
    int32_t* scan_tmp;     // scan block sums
    int32_t* counters;     // 8: [4] triangles in leaves, [5] kept inner nodes (scan totals)
    double* sah;           // 1
    FNode* fnodes;         // n (capacity)
    int2* kept_child;      // n: PLOC ids of the children of each final inner node (SOBB trees; else null)
    float4* tri_ordered;   // 3n
    int32_t* prim_order;   // n
    // SOBB builds (ndirs > 0): k-DOP of every node, 2*ndirs floats (lo..., hi...)
    int32_t ndirs;
    int32_t ntriples;
    float* kdop;           // (2n-1) x 2*ndirs
    const float* dirs;     // ndirs x 3 unit normals
    const int4* triples;   // (a, b, c, bits of 1/|det|) per non-coplanar direction triple
};

hipError_t LaunchBuildInit(const BuildBuffers& b, int block, hipStream_t s);
constexpr int kBoundsBlocks = 2048;
constexpr int kSahBlocks = 512;  // SAH partials reuse the bounds partial rows (kSahBlocks doubles fit)
hipError_t LaunchBounds(const BuildBuffers& b, hipStream_t s);
hipError_t LaunchCodes(const BuildBuffers& b, int morton_type, hipStream_t s);
hipError_t LaunchInitClusters(const BuildBuffers& b, hipStream_t s);
// lanes: 1, or 4 with block 256 (1024-thread workgroups splitting each nearest-neighbour
// search over four lanes; for iterations with about one tile per CU)
hipError_t LaunchPlocIteration(const BuildBuffers& b, int iter, int radius, int block, int lanes, int tail,
                               int max_leaf, int grid, int min_tiles, hipStream_t s);
// The same iteration as KPlocNeighbours + KPlocMerge (no workgroup waits on another), for
// iterations with several rounds of tiles per workgroup; grid: workgroups of each launch.
hipError_t LaunchPlocIterationSplit(const BuildBuffers& b, int iter, int radius, int block, int tail, int max_leaf,
                                    int grid, hipStream_t s);
hipError_t LaunchPlocTail(const BuildBuffers& b, int iter, int radius, int max_leaf, hipStream_t s);
int PlocChunk(int block);
int64_t PlocStatusWords(int64_t n, int block);  // look-back states per parity
int PlocTail();
hipError_t LaunchLayout(const BuildBuffers& b, int max_leaf, hipStream_t s);
hipError_t LaunchFitSobb(const BuildBuffers& b, hipStream_t s);
hipError_t LaunchFinish(const BuildBuffers& b, hipStream_t s);
// Two in-place exclusive scans (ny <= nx) in one set of three launches; tmp holds
// 2 ScanTempInts(nx) ints.
hipError_t LaunchExclusiveScan2(int32_t* x, int64_t nx, int32_t* y, int64_t ny, int32_t* tmp, int32_t* total_x,
                                int32_t* total_y, hipStream_t s);
size_t ScanTempInts(int64_t n);

hipError_t LaunchRadixSort(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, int64_t n,
                           int key_bits, int digit_bits, bool key64, void* temp, size_t temp_bytes, bool* result_in_alt,
                           hipStream_t s);
size_t RadixSortTempBytes(int64_t n, int key_bits, int digit_bits, bool key64);

// A built tree as the traversal kernels see it.
struct TraceTree {
    const FNode* nodes;
    const float4* tris;  // 3 float4 per primitive, leaf order
    int32_t node_count;
    int32_t prim_count;
    int32_t height;
    int32_t ndirs;        // SOBB trees: directions of the set (0 for AABB trees)
    const float* dirs;    // SOBB trees: ndirs x 3 unit normals (device)
};

// ---- batched build: many meshes, each built by one workgroup of one persistent launch ----
struct MeshResult {
    int32_t kept;         // kept inner nodes (0: the mesh is one leaf)
    int32_t leaves;
    int32_t root_height;
    int32_t iterations;   // PLOC iterations
    uint32_t flags;       // bit 0: non-finite coordinates; bit 1: an iteration without a merge
    int32_t pad;
    double sah;           // sum of the final nodes' SAH terms (before the root normalisation)
    float root_lo[3], root_hi[3];
};

struct BatchArgs {
    const float* tri;          // 9 floats per triangle, meshes back to back
    const int64_t* tri_off;    // mesh_count + 1: first triangle of each mesh
    const int64_t* fnode_off;  // mesh_count + 1: first final node (max(1, n - 1) per mesh)
    const int32_t* order;      // mesh ids by decreasing triangle count (the work queue)
    int32_t mesh_count;
    uint32_t* queue;           // next job (zero before the launch)
    // per triangle
    uint64_t* keys;
    uint64_t* keys_alt;
    uint32_t* vals;
    uint32_t* vals_alt;
    PNode* clusters0;
    PNode* clusters1;
    int32_t* leafslot;
    int32_t* keptflag;
    float4* tri_ordered;
    int32_t* prim_order;
    // per PLOC node: mesh m's 2 n_m - 1 nodes start at 2 tri_off[m] - m
    PNode* nodes;
    int32_t* parent;
    EvalRec* eval;
    uint8_t* state;
    FNode* fnodes;
    MeshResult* results;       // per mesh (zeroed before the launch)
    int morton_type, radius, max_leaf, digit_bits;
};

hipError_t LaunchBatchBuild(const BatchArgs& a, int block, int grid, hipStream_t s);
int BatchWgPerCu(int block);

// Per-mesh tree of a batched BVH, for the batched trace.
struct BatchTree {
    int64_t node_off;  // first final node
    int64_t tri_off;   // first ordered triangle
    int32_t node_count;
    int32_t prim_count;
    int32_t height;
    int32_t pad;
};

constexpr int kDplWhittedMaxDepth = 8;
hipError_t LaunchWhitted(const TraceTree& tree, const dpl_whitted_params& wp, const float* basis, const float* tri,
                         const float* vn, const int32_t* mat_id, const dpl_material* mats, const int32_t* lights,
                         int32_t light_count, float* image, int cu_count, uint32_t* counter, int32_t* spill,
                         unsigned long long* totals, hipStream_t s);

constexpr int kMaxDirs = 13;
constexpr int kMaxTriples = 286;
// The compile-time triple table of the 13-direction (26-DOP) set, for a host-side check.
int SobbTriples13();
void SobbTriple13(int t, int* a, int* b, int* c, float* inv_det);

// counters (nullable, device): accumulates {inner nodes visited, triangle tests}.
// max_grid > 0 caps the persistent grid (the spill region then needs max_grid workgroups'
// stacks); TraceGridFor gives the largest grid whose stacks fit spill_bytes.
// order (optional, n entries): the refill kernels take ray order[q] as the q-th ray to trace
// (hits still land at the ray's own index); the counter kernels ignore it.
hipError_t LaunchTrace(const TraceTree& tree, const dpl_ray* rays, int64_t n, uint32_t flags, dpl_hit* hits,
                       int cu_count, uint32_t* counter, int32_t* spill, unsigned long long* counters, hipStream_t s,
                       int max_grid = 0, const int32_t* order = nullptr);
// A coherent tracing order for n rays: Morton codes (10 bits per axis over the box lo..hi) of
// the point a quarter of the way along each ray's span inside the box, sorted with the ray
// indices (rays missing the box last). buf holds 4 n words; temp RayOrderTempBytes(n). *order
// is set to the sorted indices inside buf.
size_t RayOrderTempBytes(int64_t n);
// Ray shares of the refill traces: the rays are split into `shares` contiguous ranges (8 by
// default, one per XCD; DPL_TRACE_SHARES, 1-32), range s first taken by the workgroups b with
// b % shares == s, each with its own counter on its own 128-byte line (atomics from several
// XCDs to different words of one line serialise as badly as to one word).
constexpr int kTraceShares = 32;       // most shares
constexpr int kTraceShareStride = 32;  // words between share counters
constexpr size_t kTraceCounterBytes = size_t(kTraceShares) * kTraceShareStride * 4;
inline int TraceSharesEnv() {
    static const int v = [] {
        const char* e = getenv("DPL_TRACE_SHARES");
        const int x = e ? atoi(e) : 8;
        return x < 1 ? 1 : (x > kTraceShares ? kTraceShares : x);
    }();
    return v;
}
hipError_t LaunchRayOrder(const dpl_ray* rays, int64_t n, const float lo[3], const float hi[3], uint32_t* buf,
                          void* temp, size_t temp_bytes, const int32_t** order, hipStream_t s);
int TraceGridFor(int cu_count, int64_t n, int height, size_t spill_bytes);
// Idle lanes per wave before a persistent trace wave fetches new rays (DPL_TRACE_REFILL, 1-64;
// 0 selects the kernels that fetch 64 rays per wave once all 64 are done).
int TraceRefillLanes();
// Batched trace: ray i against mesh ray_mesh[i] of a batched BVH (per-mesh trees in mt; a
// mesh id out of range gives a miss); max_height sizes the spill stacks.
hipError_t LaunchTraceBatched(const FNode* nodes, const float4* tris, const BatchTree* mt, int32_t mesh_count,
                              int32_t max_height, const dpl_ray* rays, const int32_t* ray_mesh, int64_t n,
                              uint32_t flags, dpl_hit* hits, int cu_count, uint32_t* counter, int32_t* spill,
                              hipStream_t s);
// Bytes of the per-lane stack spill region the trace and render kernels need for a tree.
size_t TraceSpillBytes(int cu_count, int height);
// ---- wide trees: 4 or 8 children per node, 8-bit quantized child boxes ----
// Node i (of node_count, the two-child tree's count) is node i of the two-child tree opened
// greedily to W children. Bytes: origin xyz (f32), meta (biased exponent byte per axis,
// child count << 24), W child references (as FNode's), 6 planes of W bytes (lo x, hi x, lo y,
// hi y, lo z, hi z; byte k = child k); 64 B for W = 4 (zero tail), 96 B for W = 8. Child k's
// bound is fmaf(q, 2^(e - 127), origin) and contains the exact child box.
constexpr int WideNodeBytes(int w) { return w == 4 ? 64 : 96; }
struct WideTree {
    const float4* nodes;
    const float4* tris;
    int32_t node_count;
    int32_t prim_count;
    int32_t width;
    int32_t height;  // of the two-child tree (sizes the spill stacks)
};
// Axis-aligned copy of a final tree (child boxes refitted from the triangles), for SOBB trees;
// height: the tree's height (height + 1 launches). Scratch: box 6 floats and done 1 int per node.
hipError_t LaunchRefitAabb(const FNode* nodes, int32_t node_count, const float4* tris, int32_t prim_count,
                           int32_t height, FNode* out, float* box, int32_t* done, hipStream_t s);
// mt / mesh_count: the per-mesh trees of a batched BVH (node_count = all slots), else null / 0.
hipError_t LaunchWiden(const FNode* aabb_nodes, int32_t node_count, int width, void* out, const BatchTree* mt,
                       int32_t mesh_count, hipStream_t s);
hipError_t LaunchTraceWide(const WideTree& tree, const dpl_ray* rays, int64_t n, uint32_t flags, dpl_hit* hits,
                           int cu_count, uint32_t* counter, int32_t* spill, unsigned long long* counters,
                           hipStream_t s, const int32_t* order = nullptr);
size_t TraceWideSpillBytes(int cu_count, int width, int height);
// Batched BVH: ray i against mesh ray_mesh[i]; tree.nodes / tree.tris are the bases, tree.height
// the largest mesh height.
hipError_t LaunchTraceWideBatched(const WideTree& tree, const BatchTree* mt, int32_t mesh_count, const dpl_ray* rays,
                                  const int32_t* ray_mesh, int64_t n, uint32_t flags, dpl_hit* hits, int cu_count,
                                  uint32_t* counter, int32_t* spill, hipStream_t s);
hipError_t LaunchRender(bool adjoint, const TraceTree& tree, const dpl_render_params& rp, const float* basis,
                        const int32_t* mat_id, const float* albedo, const float* emission, const float* d_image,
                        float* image, float* grad_albedo, float* grad_emission, int cu_count, uint32_t* counter,
                        int32_t* spill, hipStream_t s);

}  // namespace dpl
