// C ABI of libdpl (include/dpl/dpl.h): contexts, scratch pools, build orchestration,
// trace, downloads. No exception crosses the boundary: every entry point catches and turns
// failures into dpl_status plus a message kept on the context.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "dpl_internal.h"

namespace dpl {

hipError_t LaunchSah(const FNode* nodes, const int32_t* counters, double* part, const float* sobb_dirs,
                     hipStream_t s);

// SOBB direction set (same construction as the oracle): the 13 directions of the 26-DOP or a
// subset chosen by kdop_size, and every triple with |det| >= 1e-3 in lexicographic order
// with 1/|det| computed in FP64 and rounded once.
struct SobbSetHost {
    int ndirs = 0, ntriples = 0;
    float dirs[3 * kMaxDirs];
    int4 triples[kMaxTriples];
};

SobbSetHost MakeSobbSetHost(int kdop_size) {
    const float s2 = 0.70710677f, s3 = 0.57735026f;
    const float all[13][3] = {{1, 0, 0},      {0, 1, 0},      {0, 0, 1},       {s2, s2, 0},     {s2, -s2, 0},
                              {s2, 0, s2},    {s2, 0, -s2},   {0, s2, s2},     {0, s2, -s2},    {s3, s3, s3},
                              {s3, s3, -s3},  {s3, -s3, s3},  {-s3, s3, s3}};
    std::vector<int> pick;
    if (kdop_size >= 26) {
        for (int k = 0; k < 13; k++) pick.push_back(k);
    } else if (kdop_size >= 18) {
        for (int k = 0; k < 9; k++) pick.push_back(k);
    } else if (kdop_size >= 14) {
        pick = {0, 1, 2, 9, 10, 11, 12};
    } else {
        pick = {0, 1, 2};
    }
    SobbSetHost S;
    S.ndirs = int(pick.size());
    for (int k = 0; k < S.ndirs; k++)
        for (int a = 0; a < 3; a++) S.dirs[3 * k + a] = all[pick[k]][a];
    for (int a = 0; a < S.ndirs; a++)
        for (int b = a + 1; b < S.ndirs; b++)
            for (int c = b + 1; c < S.ndirs; c++) {
                const double n1[3] = {S.dirs[3 * a], S.dirs[3 * a + 1], S.dirs[3 * a + 2]};
                const double n2[3] = {S.dirs[3 * b], S.dirs[3 * b + 1], S.dirs[3 * b + 2]};
                const double n3[3] = {S.dirs[3 * c], S.dirs[3 * c + 1], S.dirs[3 * c + 2]};
                const double cx = n2[1] * n3[2] - n2[2] * n3[1];
                const double cy = n2[2] * n3[0] - n2[0] * n3[2];
                const double cz = n2[0] * n3[1] - n2[1] * n3[0];
                const double det = n1[0] * cx + n1[1] * cy + n1[2] * cz;
                if (std::fabs(det) < 1e-3) continue;
                const float inv = float(1.0 / std::fabs(det));
                int32_t bits;
                std::memcpy(&bits, &inv, 4);
                S.triples[S.ntriples++] = make_int4(a, b, c, bits);
            }
    return S;
}

// Camera basis of src/renderer/cpu_raytracer.cc:62-73 (pos, p00, qw, qh), evaluated once on
// the host in FP32 with the same operation order.
void CameraBasisHost(const float* cam, int w, int h, float* out) {
    auto normalize = [](float x, float y, float z, float* o) {
        const float len = std::sqrt(x * x + y * y + z * z);
        if (len <= 1e-6f) {
            o[0] = o[1] = o[2] = 0.0f;
            return;
        }
        o[0] = x / len, o[1] = y / len, o[2] = z / len;
    };
    const float *dir = cam + 3, *up = cam + 6;
    const float fov = cam[9];
    float right[3];
    normalize(dir[1] * up[2] - dir[2] * up[1], dir[2] * up[0] - dir[0] * up[2], dir[0] * up[1] - dir[1] * up[0], right);
    const float gw = static_cast<float>(2.0 * std::tan(fov / 2.0f));
    const float gh = gw * (static_cast<float>(h) / static_cast<float>(w));
    const float sw = gw / static_cast<float>(w - 1), sh = gh / static_cast<float>(h - 1);
    const float hw = gw / 2, hh = gh / 2;
    for (int a = 0; a < 3; a++) {
        out[a] = cam[a];
        out[3 + a] = dir[a] - right[a] * hw + up[a] * hh;
        out[6 + a] = right[a] * sw;
        out[9 + a] = -up[a] * sh;
    }
}

struct Error : std::runtime_error {
    dpl_status status;
    Error(dpl_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define DPL_HIP(expr)                                                                                       \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess)                                                                               \
            throw ::dpl::Error(e_ == hipErrorOutOfMemory ? DPL_ERR_OOM : DPL_ERR_DEVICE,                    \
                               std::string(#expr) + ": " + hipGetErrorString(e_));                          \
    } while (0)

// Grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <typename T = void>
    T* ensure(size_t bytes) {
        bytes = std::max<size_t>(bytes, 256);
        if (bytes > cap) {
            release();
            hipError_t e = hipMalloc(&p, bytes);
            if (e != hipSuccess) {
                p = nullptr;
                (void)hipGetLastError();
                throw Error(DPL_ERR_OOM, "hipMalloc(" + std::to_string(bytes) + " bytes) failed");
            }
            cap = bytes;
        }
        return static_cast<T*>(p);
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

}  // namespace dpl

using namespace dpl;

// Scratch of dpl_trace_noalloc: the ray counter (first 256 bytes) and the traversal spill
// stacks, one copy per device for the life of the process. Host code uses only the device
// address that hipGetSymbolAddress gives for the current device (kept per device below), never
// the array's name.
namespace {
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr size_t kNoAllocCounterBytes = kTraceCounterBytes;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];
std::mutex g_scratch_mu;
std::vector<uint8_t*> g_scratch_addr;  // per device, null until looked up

// Called with `device` current, from dpl_context_create.
uint8_t* ScratchAddress(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (size_t(device) >= g_scratch_addr.size()) g_scratch_addr.resize(size_t(device) + 1, nullptr);
    if (!g_scratch_addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            throw Error(DPL_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        g_scratch_addr[device] = static_cast<uint8_t*>(p);
    }
    return g_scratch_addr[device];
}
}  // namespace

struct dpl_context {
    int device = 0;
    int cu_count = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string last_error;
    // scratch
    DevBuf tri, keys, vals, keys_alt, vals_alt, nodes, parent, clusters0, clusters1, ccount, status0, status1,
        small, eval, state, kept_id, leafslot, keptflag, scan_tmp, sort_tmp, rays, hits,
        trace_counter, spill, kdop, sobb_set, wide_tmp, ray_order, ray_order_tmp;
    // batched builds and traces
    DevBuf b_off, b_foff, b_order, b_results, b_queue, ray_mesh;
    std::string device_name;
    uint8_t* noalloc_scratch = nullptr;  // this device's address of the no-allocation scratch
    int sobb_kdop = -1;  // kdop_size the uploaded direction set was made for
    void* pinned = nullptr;          // host staging for small results (device-visible)
    int32_t* pinned_dev = nullptr;   // device address of `pinned`
    hipEvent_t ev[11] = {};
    std::vector<hipEvent_t> iter_ev;  // per-launch timing of PLOC iterations (on request)
    // Live BVH handles of this context and the generation of their device buffers. A build,
    // rebuild, batched build or widen gives the handle a new generation (from a context-wide
    // sequence, so a handle later allocated at a freed handle's address never matches it);
    // destroy removes it. A captured trace graph replays only while its BVH keeps the
    // generation it was captured with.
    uint64_t gen_seq = 0;
    std::unordered_map<const dpl_bvh*, uint64_t> bvh_gen;
    // Live trace graphs. Destroying the context detaches them and the live BVHs (their ctx
    // becomes null), so that handles destroyed after their context free only their own
    // device memory.
    std::unordered_set<dpl_trace_graph*> graphs;
    hipEvent_t IterEvent(size_t k) {
        while (iter_ev.size() <= k) {
            hipEvent_t e;
            DPL_HIP(hipEventCreate(&e));
            iter_ev.push_back(e);
        }
        return iter_ev[k];
    }
};

struct dpl_bvh {
    dpl_context* ctx = nullptr;
    int64_t prim_count = 0;
    int32_t node_count = 0;
    int32_t tree_height = 0;
    int32_t mesh_count = 1;
    dpl_build_params params{};
    dpl_build_stats stats{};
    DevBuf fnodes, tri_ordered, prim_order;
    DevBuf sobb_dirs;  // SOBB trees: the direction set's unit normals
    int32_t ndirs = 0;  // 0 for AABB trees
    int sobb_kdop = -1;
    // intermediates (keep_intermediates)
    DevBuf codes, sorted_idx, pnodes;
    std::vector<int32_t> iter_counts;
    bool has_intermediates = false;
    bool codes64 = false;
    // batched BVHs: one tree per mesh
    bool batched = false;
    DevBuf batch_tree;                 // BatchTree per mesh
    std::vector<dpl_mesh_info> meshes;
    int64_t fnode_slots = 0;           // final-node slots of all meshes
    // wide form (dpl_bvh_widen): node i converts final node i; dropped by a rebuild
    DevBuf wnodes;
    int32_t wide_width = 0;
    double widen_ms = 0.0;
    float root_lo[3] = {}, root_hi[3] = {};  // scene box (single-scene BVHs)
    dpl_bvh() = default;
    dpl_bvh(const dpl_bvh&) = delete;
    dpl_bvh& operator=(const dpl_bvh&) = delete;
    ~dpl_bvh() {
        if (ctx) ctx->bvh_gen.erase(this);  // also when a failed build drops the new handle
    }
};

namespace {
// New device buffers (or new contents) behind the handle: invalidates captured graphs.
void BumpGeneration(dpl_bvh* bvh) { bvh->ctx->bvh_gen[bvh] = ++bvh->ctx->gen_seq; }
}  // namespace

namespace {

dpl_status Fail(dpl_context* ctx, dpl_status s, const std::string& m) {
    if (ctx) ctx->last_error = m;
    return s;
}

template <typename F>
dpl_status Guard(dpl_context* ctx, F&& f) {
    try {
        if (ctx) {
            DPL_HIP(hipSetDevice(ctx->device));
        }
        f();
        if (ctx) ctx->last_error.clear();
        return DPL_OK;
    } catch (const Error& e) {
        return Fail(ctx, e.status, e.what());
    } catch (const std::bad_alloc&) {
        return Fail(ctx, DPL_ERR_OOM, "host allocation failed");
    } catch (const std::exception& e) {
        return Fail(ctx, DPL_ERR_INTERNAL, e.what());
    } catch (...) {
        return Fail(ctx, DPL_ERR_INTERNAL, "unknown error");
    }
}

void Require(bool cond, const std::string& msg) {
    if (!cond) throw Error(DPL_ERR_INVALID_ARG, msg);
}

dpl_build_params Normalize(const dpl_build_params* p) {
    dpl_build_params q = *p;
    if (q.block_size == 0) q.block_size = 256;
    if (q.sort_digit_bits == 0) q.sort_digit_bits = 8;
    if (q.kdop_size == 0) q.kdop_size = 32;
    Require(q.morton_type == DPL_MORTON30 || q.morton_type == DPL_EMC60, "morton_type must be 0 or 1");
    Require(q.bv_type == DPL_BV_AABB || q.bv_type == DPL_BV_SOBB, "bv_type must be 0 or 1");
    Require(q.nn_search_radius >= 1 && q.nn_search_radius <= kMaxRadius, "nn_search_radius must be in [1, 32]");
    Require(q.max_triangles_per_leaf >= 1 && q.max_triangles_per_leaf <= kMaxLeaf,
            "max_triangles_per_leaf must be in [1, 15]");
    Require(q.block_size == 128 || q.block_size == 256 || q.block_size == 512, "block_size must be 128, 256 or 512");
    Require(q.sort_digit_bits == 4 || q.sort_digit_bits == 8, "sort_digit_bits must be 4 or 8");
    Require(q.kdop_size >= 6, "kdop_size must be at least 6");
    return q;
}

float ElapsedMs(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0.0f;
    }
    return ms;
}


// PLOC iterations with more two-launch tiles (2 x block clusters) than this many rounds of
// the fused kernel's workgroups run as two launches (neighbours, then merges) instead of one:
// for 256-thread blocks, above ~98k clusters (2M build: 0.1 / 0.15 / 0.2 / 0.4 / 1 round
// measured 2.28 / 2.27 / 2.26 / 2.27 / 2.32 ms)
constexpr double kPlocSplitRounds = 0.2;

void BuildImpl(dpl_context* ctx, dpl_bvh* bvh, const float* tri_in, int64_t n, dpl_mem mem,
               const dpl_build_params& p, dpl_build_stats* out_stats) {
    Require(tri_in != nullptr, "tri_xyz is null");
    Require(n >= 1, "at least one triangle is required");
    Require(n < (int64_t(1) << 27), "at most 2^27 - 1 triangles per BVH");
    Require(mem == DPL_MEM_HOST || mem == DPL_MEM_DEVICE, "mem must be DPL_MEM_HOST or DPL_MEM_DEVICE");
    BumpGeneration(bvh);
    hipStream_t s = ctx->stream;
    dpl_context& C = *ctx;
    const int64_t nn = 2 * n - 1;
    const int64_t status_words = PlocStatusWords(n, p.block_size);
    const bool key64 = p.morton_type == DPL_EMC60;
    const int key_bits = key64 ? 60 : 30;

    // ---- allocate everything before the first timed event ----
    const float* tri = tri_in;
    if (mem == DPL_MEM_HOST) tri = C.tri.ensure<float>(size_t(n) * 36);
    BuildBuffers b{};
    b.n = n;
    b.tri = tri;
    b.keys = C.keys.ensure<uint64_t>(size_t(n) * 8);
    b.vals = C.vals.ensure<uint32_t>(size_t(n) * 4);
    b.keys_alt = C.keys_alt.ensure<uint64_t>(size_t(n) * 8);
    b.vals_alt = C.vals_alt.ensure<uint32_t>(size_t(n) * 4);
    b.nodes = C.nodes.ensure<PNode>(size_t(nn) * sizeof(PNode));
    b.parent = C.parent.ensure<int32_t>(size_t(nn) * 4);
    b.clusters[0] = C.clusters0.ensure<PNode>(size_t(n) * sizeof(PNode));
    b.clusters[1] = C.clusters1.ensure<PNode>(size_t(n) * sizeof(PNode));
    b.ccount = C.ccount.ensure<int32_t>(size_t(n + 2) * 4);
    b.status[0] = C.status0.ensure<uint64_t>(size_t(status_words) * 8);
    b.status[1] = C.status1.ensure<uint64_t>(size_t(status_words) * 8);
    {
        uint8_t* sm = C.small.ensure<uint8_t>(1024 + size_t(kBoundsBlocks) * 64);
        b.bounds = reinterpret_cast<uint32_t*>(sm);              // 13 words
        b.bounds_part = reinterpret_cast<uint32_t*>(sm + 1024);  // kBoundsBlocks x 16
        b.tile_counter = reinterpret_cast<uint32_t*>(sm + 64);   // 2
        b.iter_done = reinterpret_cast<int32_t*>(sm + 96);       // 1
        b.counters = reinterpret_cast<int32_t*>(sm + 128);       // 8
        b.sah = reinterpret_cast<double*>(sm + 192);             // 1
    }
    b.eval = C.eval.ensure<EvalRec>(size_t(nn) * sizeof(EvalRec));
    b.state = C.state.ensure<uint8_t>(size_t(nn));
    b.results = reinterpret_cast<BuildResults*>(C.pinned_dev);
    b.leafslot = C.leafslot.ensure<int32_t>(size_t(n) * 4);
    b.keptflag = C.keptflag.ensure<int32_t>(size_t(n) * 4);
    b.scan_tmp = C.scan_tmp.ensure<int32_t>(2 * ScanTempInts(n) * 4);  // the layout scans two arrays at once
    const size_t sort_bytes = RadixSortTempBytes(n, key_bits, p.sort_digit_bits, key64);
    void* sort_tmp = C.sort_tmp.ensure(sort_bytes);
    b.fnodes = bvh->fnodes.ensure<FNode>(size_t(std::max<int64_t>(1, n - 1)) * sizeof(FNode));
    b.kept_child = p.bv_type == DPL_BV_SOBB ? C.kept_id.ensure<int2>(size_t(n) * 8) : nullptr;
    b.tri_ordered = bvh->tri_ordered.ensure<float4>(size_t(n) * 48);
    b.prim_order = bvh->prim_order.ensure<int32_t>(size_t(n) * 4);
    if (p.bv_type == DPL_BV_SOBB) {
        const SobbSetHost S = MakeSobbSetHost(p.kdop_size);
        uint8_t* sd = C.sobb_set.ensure<uint8_t>(sizeof(S.dirs) + sizeof(S.triples));
        if (C.sobb_kdop != p.kdop_size) {
            DPL_HIP(hipMemcpy(sd, S.dirs, sizeof(S.dirs), hipMemcpyHostToDevice));
            DPL_HIP(hipMemcpy(sd + sizeof(S.dirs), S.triples, sizeof(S.triples), hipMemcpyHostToDevice));
            C.sobb_kdop = p.kdop_size;
        }
        b.ndirs = S.ndirs;
        b.ntriples = S.ntriples;
        b.dirs = reinterpret_cast<const float*>(sd);
        b.triples = reinterpret_cast<const int4*>(sd + sizeof(S.dirs));
        if (S.ndirs == 13) {  // the 26-DOP fit enumerates its triples at compile time
            bool same = S.ntriples == SobbTriples13();
            for (int t = 0; same && t < S.ntriples; t++) {
                int a, c, d;
                float inv;
                SobbTriple13(t, &a, &c, &d, &inv);
                int32_t bits;
                std::memcpy(&bits, &inv, 4);
                same = a == S.triples[t].x && c == S.triples[t].y && d == S.triples[t].z && bits == S.triples[t].w;
            }
            if (!same) throw Error(DPL_ERR_INTERNAL, "26-DOP triple table mismatch");
        }
        // non-axis slabs only (the axis slabs are the node box -/+ pad)
        b.kdop = C.kdop.ensure<float>(std::max<size_t>(16, size_t(nn) * 2 * (S.ndirs - 3) * 4));
        if (bvh->sobb_kdop != p.kdop_size) {
            float* bd = bvh->sobb_dirs.ensure<float>(sizeof(S.dirs));
            DPL_HIP(hipMemcpy(bd, S.dirs, sizeof(S.dirs), hipMemcpyHostToDevice));
            bvh->sobb_kdop = p.kdop_size;
        }
    }
    if (p.keep_intermediates) {
        bvh->codes.ensure(size_t(n) * 8);
        bvh->sorted_idx.ensure(size_t(n) * 4);
        bvh->pnodes.ensure(size_t(nn) * sizeof(PNode));
    }

    dpl_build_stats st{};
    if (mem == DPL_MEM_HOST) {
        DPL_HIP(hipEventRecord(C.ev[8], s));
        DPL_HIP(hipMemcpyAsync(const_cast<float*>(tri), tri_in, size_t(n) * 36, hipMemcpyHostToDevice, s));
        DPL_HIP(hipEventRecord(C.ev[9], s));
    }

    // per-stage and per-launch events only on request (keep_intermediates bit 1): an event
    // recorded between two kernels leaves the GPU idle for about 5 us
    const bool time_iters = (p.keep_intermediates & 2) != 0;
    auto stage_event = [&](int k) {
        if (time_iters) DPL_HIP(hipEventRecord(C.ev[k], s));
    };
    DPL_HIP(hipEventRecord(C.ev[0], s));
    DPL_HIP(LaunchBuildInit(b, p.block_size, s));
    DPL_HIP(LaunchBounds(b, s));
    stage_event(1);
    DPL_HIP(LaunchCodes(b, p.morton_type, s));
    if (p.keep_intermediates) {
        DPL_HIP(hipMemcpyAsync(bvh->codes.p, b.keys, size_t(n) * (key64 ? 8 : 4), hipMemcpyDeviceToDevice, s));
    }
    stage_event(2);
    bool in_alt = false;
    DPL_HIP(LaunchRadixSort(b.keys, b.vals, b.keys_alt, b.vals_alt, n, key_bits, p.sort_digit_bits, key64, sort_tmp,
                            sort_bytes, &in_alt, s));
    if (in_alt) {
        std::swap(b.keys, b.keys_alt);
        std::swap(b.vals, b.vals_alt);
    }
    b.pairs = b.vals_alt;  // free once the sort is done
    stage_event(3);
    DPL_HIP(LaunchInitClusters(b, s));

    // ---- PLOC iterations: multi-workgroup kernels while clusters exceed the tail size.
    // The host launches a batch of iterations sized from the observed reduction rate, then
    // reads the cluster count the last finished iteration wrote to host-mapped memory;
    // iterations launched after the count fell below the tail size pass the clusters through.
    volatile BuildResults* pr = static_cast<volatile BuildResults*>(C.pinned);
    pr->progress = int32_t(n);
    pr->flags = 0u;
    pr->multi_iters = 0;
    pr->tail_iters = 0;
    const int tail = PlocTail();
    int iter = 0;
    int64_t c_known = n;
    double rate = 1.15;  // clusters shrink by about this factor per iteration
    // workgroups per launch, also the tile count below which an iteration takes one cluster per
    // thread (2 above it): 4 per CU for 256 threads (PLOC iterations 3 % slower at 3 or 8 per CU,
    // 12 % at 2, 13 % at 16)
    const int cap = std::max(1, C.cu_count * (p.block_size == 512 ? 4 : (p.block_size == 128 ? 16 : 4)));
    while (c_known > tail) {
        int est = int(std::ceil(std::log(double(c_known) / tail) / std::log(rate)));
        est = std::max(1, std::min(est, 96));
        const int64_t T = (c_known + p.block_size - 1) / p.block_size;  // tiles at one cluster per thread
        const int64_t per_tile = p.block_size - 2 * p.nn_search_radius;
        // tiles of the two-launch iteration (2 x block clusters) at the known count: an upper
        // bound for every iteration of this batch
        const int64_t split_chunk = PlocChunk(p.block_size) - 2 * p.nn_search_radius;
        const int64_t split_tiles = (c_known + split_chunk - 1) / split_chunk;
        for (int g = 0; g < est; g++) {
            // once the modelled count (c_known / rate^g) needs at most two tiles per CU (two
            // 1024-thread workgroups are resident per CU): four lanes per neighbour search, twice
            // the modelled tiles up to 2 per CU (surplus workgroups leave at once; more tiles
            // than workgroups take further rounds)
            const int64_t t4 = int64_t(std::ceil(double(c_known) / std::pow(rate, g) / double(per_tile)));
            const bool four = p.block_size == 256 && t4 <= 2 * int64_t(C.cu_count);
            const int grid = int(four ? std::min<int64_t>(2 * t4, 2 * int64_t(C.cu_count)) : std::min<int64_t>(T, cap));
            // more two-launch tiles than kPlocSplitRounds rounds of the fused kernel's
            // workgroups: two launches that never wait on each other
            const int64_t t2 = int64_t(std::ceil(double(c_known) / std::pow(rate, g) / double(split_chunk)));
            const bool split = double(t2) > kPlocSplitRounds * double(cap);
            if (time_iters) DPL_HIP(hipEventRecord(C.IterEvent(2 * size_t(iter)), s));
            if (split)
                DPL_HIP(LaunchPlocIterationSplit(b, iter, p.nn_search_radius, p.block_size, tail,
                                                 p.max_triangles_per_leaf, int(std::min<int64_t>(split_tiles, 1 << 16)), s));
            else
                DPL_HIP(LaunchPlocIteration(b, iter, p.nn_search_radius, p.block_size, four ? 4 : 1, tail,
                                            p.max_triangles_per_leaf, grid, cap, s));
            if (time_iters) DPL_HIP(hipEventRecord(C.IterEvent(2 * size_t(iter) + 1), s));
            iter++;
        }
        DPL_HIP(hipStreamSynchronize(s));
        if (pr->flags & 1u) throw Error(DPL_ERR_INVALID_ARG, "triangle coordinates must be finite");
        const int64_t c_now = pr->progress;
        if (c_now >= c_known) throw Error(DPL_ERR_INTERNAL, "PLOC made no progress");
        rate = std::max(1.02, std::min(2.0, std::pow(double(c_known) / double(c_now), 1.0 / est)));
        c_known = c_now;
    }
    DPL_HIP(LaunchPlocTail(b, iter, p.nn_search_radius, p.max_triangles_per_leaf, s));
    stage_event(4);
    stage_event(5);
    DPL_HIP(LaunchLayout(b, p.max_triangles_per_leaf, s));
    stage_event(6);
    DPL_HIP(LaunchFitSobb(b, s));
    stage_event(10);
    DPL_HIP(LaunchSah(b.fnodes, b.counters + 5, reinterpret_cast<double*>(b.bounds_part),
                      b.ndirs > 0 ? b.dirs : nullptr, s));
    DPL_HIP(LaunchFinish(b, s));
    DPL_HIP(hipEventRecord(C.ev[7], s));
    if (p.keep_intermediates) {
        DPL_HIP(hipMemcpyAsync(bvh->sorted_idx.p, b.vals, size_t(n) * 4, hipMemcpyDeviceToDevice, s));
        DPL_HIP(hipMemcpyAsync(bvh->pnodes.p, b.nodes, size_t(nn) * sizeof(PNode), hipMemcpyDeviceToDevice, s));
    }
    DPL_HIP(hipStreamSynchronize(s));
    if (pr->flags & 1u) throw Error(DPL_ERR_INVALID_ARG, "triangle coordinates must be finite");
    if (pr->flags & 2u) throw Error(DPL_ERR_INTERNAL, "PLOC iteration without a merge");

    const int32_t K = pr->kept;
    const int32_t leaves = pr->leaves;
    const int32_t root_height = pr->root_height;
    const int32_t real_iters = pr->multi_iters + pr->tail_iters;
    const double sah_sum = pr->sah;
    float lo[3], hi[3];
    for (int a = 0; a < 3; a++) lo[a] = pr->root_lo[a], hi[a] = pr->root_hi[a];
    std::vector<int32_t> iters;
    const int32_t multi = pr->multi_iters;
    if (p.keep_intermediates) {
        // clusters at the start of every iteration (pass-through launches repeat a count)
        std::vector<int32_t> cc(size_t(pr->iter_done) + 1);
        DPL_HIP(hipMemcpy(cc.data(), b.ccount, cc.size() * 4, hipMemcpyDeviceToHost));
        for (int k = 0; k < pr->iter_done; k++)
            if (iters.empty() || cc[k] != iters.back()) iters.push_back(cc[k]);
        if (!iters.empty() && iters.back() == 1) iters.pop_back();
        if (time_iters) {
            // the first `multi` launches merged; per launch: read c_k clusters (32 B), write
            // c_k+1 survivors (32 B), per merge a node (32 B), two parents (8 B), its
            // collapse record (17 B) and the two children's records (34 B); with SOBB bounds
            // also the two children's stored k-DOP slabs and the new node's (3 x 8 (ndirs - 3) B)
            const double per_merge = 91.0 + (b.ndirs > 3 ? 3.0 * 8.0 * (b.ndirs - 3) : 0.0);
            for (int k = 0; k < multi; k++) {
                st.ploc_launch_ms += ElapsedMs(C.iter_ev[2 * k], C.iter_ev[2 * k + 1]);
                const double ck = cc[k], ck1 = cc[k + 1];
                st.ploc_bytes += 32.0 * ck + 32.0 * ck1 + per_merge * (ck - ck1);
            }
        }
    }
    st.ploc_launches = multi;

    st.build_ms = ElapsedMs(C.ev[0], C.ev[7]);
    st.upload_ms = mem == DPL_MEM_HOST ? ElapsedMs(C.ev[8], C.ev[9]) : 0.0;
    if (time_iters) {
        st.stage_ms[0] = ElapsedMs(C.ev[0], C.ev[1]);
        st.stage_ms[1] = ElapsedMs(C.ev[1], C.ev[2]);
        st.stage_ms[2] = ElapsedMs(C.ev[2], C.ev[3]);
        st.stage_ms[3] = ElapsedMs(C.ev[3], C.ev[4]);
        st.stage_ms[4] = ElapsedMs(C.ev[4], C.ev[5]);
        st.stage_ms[5] = ElapsedMs(C.ev[5], C.ev[6]);
        st.stage_ms[6] = ElapsedMs(C.ev[6], C.ev[10]);
        st.stage_ms[7] = ElapsedMs(C.ev[10], C.ev[7]);
    }
    const int32_t fcount = K > 0 ? K : 1;
    st.inner_node_count = fcount;
    st.leaf_node_count = leaves;
    st.node_count = fcount + leaves;
    st.ploc_iterations = real_iters;
    st.memory_consumption = int64_t(fcount) * 64 + n * 48 + n * 4;
    const double sx = double(hi[0]) - lo[0], sy = double(hi[1]) - lo[1], sz = double(hi[2]) - lo[2];
    const double root_sa = 2.0 * (sx * sy + sy * sz + sx * sz);
    if (K > 0) {
        st.sah_cost = root_sa > 0 ? (root_sa + sah_sum) / root_sa : 0.0;
    } else {
        st.sah_cost = double(n);  // a single leaf: Ci * SA * n / SA
    }
    st.tree_height = root_height;

    bvh->prim_count = n;
    bvh->node_count = fcount;
    bvh->tree_height = root_height;
    bvh->ndirs = b.ndirs;
    for (int a = 0; a < 3; a++) bvh->root_lo[a] = lo[a], bvh->root_hi[a] = hi[a];
    bvh->params = p;
    bvh->stats = st;
    bvh->iter_counts = iters;
    bvh->has_intermediates = p.keep_intermediates != 0;
    bvh->codes64 = key64;
    bvh->wide_width = 0;
    if (out_stats) *out_stats = st;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------
extern "C" {

dpl_status dpl_build_params_from_type(int32_t accel_type, int32_t radius, int32_t max_leaf, int32_t kdop,
                                      dpl_build_params* p) {
    if (!p) return DPL_ERR_INVALID_ARG;
    dpl_build_params q{};
    switch (accel_type) {
        case DPL_ACCEL_PLOC: q.morton_type = DPL_MORTON30, q.bv_type = DPL_BV_AABB; break;
        case DPL_ACCEL_PLOC_EMC: q.morton_type = DPL_EMC60, q.bv_type = DPL_BV_AABB; break;
        case DPL_ACCEL_PLOC_SOBB: q.morton_type = DPL_MORTON30, q.bv_type = DPL_BV_SOBB; break;
        case DPL_ACCEL_PLOC_EMC_SOBB: q.morton_type = DPL_EMC60, q.bv_type = DPL_BV_SOBB; break;
        case DPL_ACCEL_DUMMY:
        case DPL_ACCEL_SBVH: return DPL_ERR_UNSUPPORTED;
        default: return DPL_ERR_INVALID_ARG;
    }
    q.nn_search_radius = radius;
    q.max_triangles_per_leaf = max_leaf;
    q.kdop_size = kdop;
    *p = q;
    return DPL_OK;
}

dpl_status dpl_context_create(int32_t device, dpl_context** out) {
    if (!out) return DPL_ERR_INVALID_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return DPL_ERR_DEVICE;
    }
    if (device < 0 || device >= count) return DPL_ERR_INVALID_ARG;
    std::unique_ptr<dpl_context> ctx(new (std::nothrow) dpl_context());
    if (!ctx) return DPL_ERR_OOM;
    ctx->device = device;
    dpl_status st = Guard(ctx.get(), [&] {
        hipDeviceProp_t prop;
        DPL_HIP(hipGetDeviceProperties(&prop, device));
        ctx->cu_count = prop.multiProcessorCount;
        ctx->device_name = std::string(prop.gcnArchName) + " " + prop.name;
        DPL_HIP(hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking));
        ctx->stream = ctx->own_stream;
        for (auto& e : ctx->ev) DPL_HIP(hipEventCreate(&e));
        DPL_HIP(hipHostMalloc(&ctx->pinned, 4096, hipHostMallocMapped));
        DPL_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&ctx->pinned_dev), ctx->pinned, 0));
        ctx->trace_counter.ensure(kTraceCounterBytes);
        ctx->noalloc_scratch = ScratchAddress(device);
    });
    if (st != DPL_OK) return st;
    *out = ctx.release();
    return DPL_OK;
}

static void DetachGraphs(dpl_context* ctx);

void dpl_context_destroy(dpl_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->bvh_gen) const_cast<dpl_bvh*>(kv.first)->ctx = nullptr;
    ctx->bvh_gen.clear();
    DetachGraphs(ctx);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : ctx->iter_ev) (void)hipEventDestroy(e);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
}

const char* dpl_last_error(const dpl_context* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

dpl_status dpl_context_set_stream(dpl_context* ctx, void* stream) {
    if (!ctx) return DPL_ERR_INVALID_ARG;
    ctx->stream = static_cast<hipStream_t>(stream);
    return DPL_OK;
}

dpl_status dpl_context_synchronize(dpl_context* ctx) {
    if (!ctx) return DPL_ERR_INVALID_ARG;
    return Guard(ctx, [&] { DPL_HIP(hipStreamSynchronize(ctx->stream)); });
}

int32_t dpl_context_cu_count(const dpl_context* ctx) { return ctx ? ctx->cu_count : 0; }

dpl_status dpl_bvh_build(dpl_context* ctx, const float* tri, int64_t n, dpl_mem mem, const dpl_build_params* params,
                         dpl_bvh** out, dpl_build_stats* stats) {
    if (!ctx || !out || !params) return ctx ? Fail(ctx, DPL_ERR_INVALID_ARG, "null argument") : DPL_ERR_INVALID_ARG;
    *out = nullptr;
    std::unique_ptr<dpl_bvh> bvh(new (std::nothrow) dpl_bvh());
    if (!bvh) return Fail(ctx, DPL_ERR_OOM, "host allocation failed");
    bvh->ctx = ctx;
    dpl_status st = Guard(ctx, [&] {
        dpl_build_params p = Normalize(params);
        BuildImpl(ctx, bvh.get(), tri, n, mem, p, stats);
    });
    if (st != DPL_OK) return st;
    *out = bvh.release();
    return DPL_OK;
}

dpl_status dpl_bvh_rebuild(dpl_bvh* bvh, const float* tri, int64_t n, dpl_mem mem, const dpl_build_params* params,
                           dpl_build_stats* stats) {
    if (!bvh || !bvh->ctx || !params) return DPL_ERR_INVALID_ARG;
    return Guard(bvh->ctx, [&] {
        dpl_build_params p = Normalize(params);
        BuildImpl(bvh->ctx, bvh, tri, n, mem, p, stats);
    });
}

namespace {

// Batched build: validates the meshes, uploads the offset tables and the work queue order
// (meshes by decreasing size), launches the persistent per-mesh build once, and reads the
// per-mesh results back to fill the mesh table, the per-mesh trace trees and the stats.
void BatchImpl(dpl_context* ctx, dpl_bvh* bvh, const float* tri_in, const int64_t* offsets, int32_t M, dpl_mem mem,
               const dpl_build_params& p, dpl_build_stats* out_stats) {
    Require(tri_in != nullptr && offsets != nullptr, "tri_xyz and mesh_offsets must be non-null");
    Require(M >= 1, "at least one mesh is required");
    Require(mem == DPL_MEM_HOST || mem == DPL_MEM_DEVICE, "mem must be DPL_MEM_HOST or DPL_MEM_DEVICE");
    Require(offsets[0] == 0, "mesh_offsets[0] must be 0");
    for (int32_t m = 0; m < M; m++) {
        Require(offsets[m + 1] > offsets[m], "every mesh needs at least one triangle (offsets strictly increasing)");
        Require(offsets[m + 1] - offsets[m] < (int64_t(1) << 27), "at most 2^27 - 1 triangles per mesh");
    }
    const int64_t N = offsets[M];
    Require(N < (int64_t(1) << 31), "at most 2^31 - 1 triangles per batch");
    if (p.bv_type != DPL_BV_AABB) throw Error(DPL_ERR_UNSUPPORTED, "batched builds support AABB bounds only");
    Require(p.block_size == 256 || p.block_size == 512, "batched block_size must be 256 or 512");
    BumpGeneration(bvh);
    hipStream_t s = ctx->stream;
    dpl_context& C = *ctx;

    std::vector<int64_t> foff(size_t(M) + 1, 0);
    for (int32_t m = 0; m < M; m++) foff[m + 1] = foff[m] + std::max<int64_t>(1, offsets[m + 1] - offsets[m] - 1);
    std::vector<int32_t> order(M);
    for (int32_t m = 0; m < M; m++) order[m] = m;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return offsets[a + 1] - offsets[a] > offsets[b + 1] - offsets[b];
    });
    const int64_t nodes_total = 2 * N - M;

    BatchArgs A{};
    const float* tri = tri_in;
    if (mem == DPL_MEM_HOST) tri = C.tri.ensure<float>(size_t(N) * 36);
    A.tri = tri;
    A.tri_off = C.b_off.ensure<int64_t>(size_t(M + 1) * 8);
    A.fnode_off = C.b_foff.ensure<int64_t>(size_t(M + 1) * 8);
    A.order = C.b_order.ensure<int32_t>(size_t(M) * 4);
    A.mesh_count = M;
    A.queue = C.b_queue.ensure<uint32_t>(64);
    A.keys = C.keys.ensure<uint64_t>(size_t(N) * 8);
    A.keys_alt = C.keys_alt.ensure<uint64_t>(size_t(N) * 8);
    A.vals = C.vals.ensure<uint32_t>(size_t(N) * 4);
    A.vals_alt = C.vals_alt.ensure<uint32_t>(size_t(N) * 4);
    A.clusters0 = C.clusters0.ensure<PNode>(size_t(N) * sizeof(PNode));
    A.clusters1 = C.clusters1.ensure<PNode>(size_t(N) * sizeof(PNode));
    A.leafslot = C.leafslot.ensure<int32_t>(size_t(N) * 4);
    A.keptflag = C.keptflag.ensure<int32_t>(size_t(N) * 4);
    A.nodes = C.nodes.ensure<PNode>(size_t(nodes_total) * sizeof(PNode));
    A.parent = C.parent.ensure<int32_t>(size_t(nodes_total) * 4);
    A.eval = C.eval.ensure<EvalRec>(size_t(nodes_total) * sizeof(EvalRec));
    A.state = C.state.ensure<uint8_t>(size_t(nodes_total));
    A.results = C.b_results.ensure<MeshResult>(size_t(M) * sizeof(MeshResult));
    A.fnodes = bvh->fnodes.ensure<FNode>(size_t(foff[M]) * sizeof(FNode));
    A.tri_ordered = bvh->tri_ordered.ensure<float4>(size_t(N) * 48);
    A.prim_order = bvh->prim_order.ensure<int32_t>(size_t(N) * 4);
    A.morton_type = p.morton_type;
    A.radius = p.nn_search_radius;
    A.max_leaf = p.max_triangles_per_leaf;
    A.digit_bits = p.sort_digit_bits;
    BatchTree* trees = bvh->batch_tree.ensure<BatchTree>(size_t(M) * sizeof(BatchTree));

    if (mem == DPL_MEM_HOST) {
        DPL_HIP(hipEventRecord(C.ev[8], s));
        DPL_HIP(hipMemcpyAsync(const_cast<float*>(A.tri), tri_in, size_t(N) * 36, hipMemcpyHostToDevice, s));
        DPL_HIP(hipEventRecord(C.ev[9], s));
    }
    DPL_HIP(hipMemcpyAsync(const_cast<int64_t*>(A.tri_off), offsets, size_t(M + 1) * 8, hipMemcpyHostToDevice, s));
    DPL_HIP(hipMemcpyAsync(const_cast<int64_t*>(A.fnode_off), foff.data(), size_t(M + 1) * 8, hipMemcpyHostToDevice,
                           s));
    DPL_HIP(hipMemcpyAsync(const_cast<int32_t*>(A.order), order.data(), size_t(M) * 4, hipMemcpyHostToDevice, s));
    DPL_HIP(hipMemsetAsync(A.results, 0, size_t(M) * sizeof(MeshResult), s));
    DPL_HIP(hipMemsetAsync(A.queue, 0, 64, s));
    const int grid = int(std::min<int64_t>(M, int64_t(C.cu_count) * BatchWgPerCu(p.block_size)));
    DPL_HIP(hipEventRecord(C.ev[0], s));
    DPL_HIP(LaunchBatchBuild(A, p.block_size, grid, s));
    DPL_HIP(hipEventRecord(C.ev[7], s));
    std::vector<MeshResult> res(M);
    DPL_HIP(hipMemcpyAsync(res.data(), A.results, size_t(M) * sizeof(MeshResult), hipMemcpyDeviceToHost, s));
    DPL_HIP(hipStreamSynchronize(s));
    for (const MeshResult& r : res) {
        if (r.flags & 1u) throw Error(DPL_ERR_INVALID_ARG, "triangle coordinates must be finite");
        if (r.flags & 2u) throw Error(DPL_ERR_INTERNAL, "PLOC iteration without a merge");
    }

    dpl_build_stats st{};
    std::vector<BatchTree> bt(M);
    bvh->meshes.assign(M, dpl_mesh_info{});
    double sah_weighted = 0.0;
    int32_t max_height = 0, max_iters = 0;
    int64_t inner = 0, leaves = 0;
    for (int32_t m = 0; m < M; m++) {
        const MeshResult& r = res[m];
        const int64_t nm = offsets[m + 1] - offsets[m];
        const int32_t fcount = r.kept > 0 ? r.kept : 1;
        const double sx = double(r.root_hi[0]) - r.root_lo[0], sy = double(r.root_hi[1]) - r.root_lo[1],
                     sz = double(r.root_hi[2]) - r.root_lo[2];
        const double root_sa = 2.0 * (sx * sy + sy * sz + sx * sz);
        const double sah = r.kept > 0 ? (root_sa > 0 ? (root_sa + r.sah) / root_sa : 0.0) : double(nm);
        dpl_mesh_info& mi = bvh->meshes[m];
        mi.node_offset = foff[m];
        mi.prim_offset = offsets[m];
        mi.node_count = fcount;
        mi.prim_count = int32_t(nm);
        mi.tree_height = r.root_height;
        mi.ploc_iterations = r.iterations;
        mi.sah_cost = sah;
        bt[m] = BatchTree{foff[m], offsets[m], fcount, int32_t(nm), r.root_height, 0};
        sah_weighted += sah * double(nm);
        max_height = std::max(max_height, r.root_height);
        max_iters = std::max(max_iters, r.iterations);
        inner += fcount;
        leaves += r.leaves;
    }
    DPL_HIP(hipMemcpyAsync(trees, bt.data(), size_t(M) * sizeof(BatchTree), hipMemcpyHostToDevice, s));
    DPL_HIP(hipStreamSynchronize(s));
    st.build_ms = ElapsedMs(C.ev[0], C.ev[7]);
    st.upload_ms = mem == DPL_MEM_HOST ? ElapsedMs(C.ev[8], C.ev[9]) : 0.0;
    st.stage_ms[3] = st.build_ms;  // one launch runs every stage
    st.inner_node_count = int32_t(std::min<int64_t>(inner, INT32_MAX));
    st.leaf_node_count = int32_t(std::min<int64_t>(leaves, INT32_MAX));
    st.node_count = int32_t(std::min<int64_t>(inner + leaves, INT32_MAX));
    st.ploc_iterations = max_iters;
    st.memory_consumption = int64_t(inner) * 64 + N * 48 + N * 4;
    st.sah_cost = sah_weighted / double(N);
    st.tree_height = max_height;

    bvh->batched = true;
    bvh->mesh_count = M;
    bvh->prim_count = N;
    bvh->node_count = int32_t(std::min<int64_t>(foff[M], INT32_MAX));
    bvh->fnode_slots = foff[M];
    bvh->tree_height = max_height;
    bvh->ndirs = 0;
    bvh->params = p;
    bvh->has_intermediates = false;
    bvh->wide_width = 0;
    bvh->stats = st;
    if (out_stats) *out_stats = st;
}

}  // namespace

dpl_status dpl_bvh_build_batched(dpl_context* ctx, const float* tri, const int64_t* offsets, int32_t mesh_count,
                                 dpl_mem mem, const dpl_build_params* params, dpl_bvh** out, dpl_build_stats* stats) {
    if (!ctx || !out || !params) return ctx ? Fail(ctx, DPL_ERR_INVALID_ARG, "null argument") : DPL_ERR_INVALID_ARG;
    *out = nullptr;
    std::unique_ptr<dpl_bvh> bvh(new (std::nothrow) dpl_bvh());
    if (!bvh) return Fail(ctx, DPL_ERR_OOM, "host allocation failed");
    bvh->ctx = ctx;
    dpl_status st = Guard(ctx, [&] {
        dpl_build_params p = Normalize(params);
        BatchImpl(ctx, bvh.get(), tri, offsets, mesh_count, mem, p, stats);
    });
    if (st != DPL_OK) return st;
    *out = bvh.release();
    return DPL_OK;
}

dpl_status dpl_bvh_get_meshes(const dpl_bvh* bvh, dpl_mesh_info* out) {
    if (!bvh || !out) return DPL_ERR_INVALID_ARG;
    if (bvh->batched) {
        std::memcpy(out, bvh->meshes.data(), bvh->meshes.size() * sizeof(dpl_mesh_info));
    } else {
        dpl_mesh_info mi{};
        mi.node_count = bvh->node_count;
        mi.prim_count = int32_t(bvh->prim_count);
        mi.tree_height = bvh->tree_height;
        mi.ploc_iterations = bvh->stats.ploc_iterations;
        mi.sah_cost = bvh->stats.sah_cost;
        out[0] = mi;
    }
    return DPL_OK;
}

void dpl_bvh_destroy(dpl_bvh* bvh) {
    if (!bvh) return;
    if (bvh->ctx) {
        (void)hipSetDevice(bvh->ctx->device);
        (void)hipStreamSynchronize(bvh->ctx->stream);
        bvh->ctx->bvh_gen.erase(bvh);
    }
    delete bvh;
}

dpl_status dpl_bvh_get_stats(const dpl_bvh* bvh, dpl_build_stats* stats) {
    if (!bvh || !stats) return DPL_ERR_INVALID_ARG;
    *stats = bvh->stats;
    return DPL_OK;
}

dpl_status dpl_bvh_sah_aabb(const dpl_bvh* bvh, double* sah) {
    if (!bvh || !bvh->ctx || !sah) return DPL_ERR_INVALID_ARG;
    dpl_context* ctx = bvh->ctx;
    return Guard(ctx, [&] {
        // AABB trees (and batched BVHs, which are AABB only): the build's own SAH; a
        // single-leaf tree costs n either way
        if (bvh->ndirs == 0 || bvh->batched || bvh->stats.leaf_node_count <= 1) {
            *sah = bvh->stats.sah_cost;
            return;
        }
        hipStream_t s = ctx->stream;
        const int32_t K = bvh->node_count;
        // refit (as for wide trees) into the context's scratch, then the same SAH kernel with
        // axis-aligned areas: partial sums per workgroup, added here in workgroup order
        const size_t refit_bytes = size_t(K) * (sizeof(FNode) + 28);
        uint8_t* t = ctx->wide_tmp.ensure<uint8_t>(refit_bytes + 256 + size_t(kSahBlocks) * 8);
        FNode* refit = reinterpret_cast<FNode*>(t);
        float* box = reinterpret_cast<float*>(t + size_t(K) * sizeof(FNode));
        int32_t* done = reinterpret_cast<int32_t*>(box + 6 * size_t(K));
        int32_t* dK = reinterpret_cast<int32_t*>(t + refit_bytes);
        double* part = reinterpret_cast<double*>(t + refit_bytes + 256);
        DPL_HIP(hipMemcpyAsync(dK, &K, 4, hipMemcpyHostToDevice, s));
        DPL_HIP(LaunchRefitAabb(bvh->fnodes.as<FNode>(), K, bvh->tri_ordered.as<float4>(), int32_t(bvh->prim_count),
                                bvh->tree_height, refit, box, done, s));
        DPL_HIP(LaunchSah(refit, dK, part, nullptr, s));
        std::vector<double> h(kSahBlocks);
        DPL_HIP(hipMemcpyAsync(h.data(), part, size_t(kSahBlocks) * 8, hipMemcpyDeviceToHost, s));
        DPL_HIP(hipStreamSynchronize(s));
        double sum = 0.0;
        for (double v : h) sum += v;
        const double sx = double(bvh->root_hi[0]) - bvh->root_lo[0], sy = double(bvh->root_hi[1]) - bvh->root_lo[1],
                     sz = double(bvh->root_hi[2]) - bvh->root_lo[2];
        const double root_sa = 2.0 * (sx * sy + sy * sz + sx * sz);
        *sah = root_sa > 0 ? (root_sa + sum) / root_sa : 0.0;
    });
}

dpl_status dpl_bvh_get_layout(const dpl_bvh* bvh, dpl_bvh_layout* l) {
    if (!bvh || !l) return DPL_ERR_INVALID_ARG;
    l->prim_count = bvh->prim_count;
    l->node_count = bvh->batched ? bvh->fnode_slots : bvh->node_count;
    l->ploc_node_count = bvh->has_intermediates ? 2 * bvh->prim_count - 1 : 0;
    l->iteration_count = bvh->has_intermediates ? int64_t(bvh->iter_counts.size()) : 0;
    l->mesh_count = bvh->mesh_count;
    l->node_bytes = 64;
    return DPL_OK;
}

dpl_status dpl_bvh_download(const dpl_bvh* bvh, void* nodes, void* tri_ordered, int32_t* prim_order, uint64_t* codes,
                            uint32_t* sorted_idx, void* ploc_nodes, int32_t* iter_counts) {
    if (!bvh || !bvh->ctx) return DPL_ERR_INVALID_ARG;
    dpl_context* ctx = bvh->ctx;
    return Guard(ctx, [&] {
        hipStream_t s = ctx->stream;
        DPL_HIP(hipStreamSynchronize(s));
        const int64_t n = bvh->prim_count;
        const int64_t node_slots = bvh->batched ? bvh->fnode_slots : bvh->node_count;
        if (nodes) DPL_HIP(hipMemcpy(nodes, bvh->fnodes.p, size_t(node_slots) * 64, hipMemcpyDeviceToHost));
        if (tri_ordered) DPL_HIP(hipMemcpy(tri_ordered, bvh->tri_ordered.p, size_t(n) * 48, hipMemcpyDeviceToHost));
        if (prim_order) DPL_HIP(hipMemcpy(prim_order, bvh->prim_order.p, size_t(n) * 4, hipMemcpyDeviceToHost));
        const bool want = codes || sorted_idx || ploc_nodes || iter_counts;
        if (want && !bvh->has_intermediates)
            throw Error(DPL_ERR_INVALID_ARG, "intermediates were not kept (build with keep_intermediates=1)");
        if (codes) {
            if (bvh->codes64) {
                DPL_HIP(hipMemcpy(codes, bvh->codes.p, size_t(n) * 8, hipMemcpyDeviceToHost));
            } else {
                std::vector<uint32_t> c32(n);
                DPL_HIP(hipMemcpy(c32.data(), bvh->codes.p, size_t(n) * 4, hipMemcpyDeviceToHost));
                for (int64_t i = 0; i < n; i++) codes[i] = c32[i];
            }
        }
        if (sorted_idx) DPL_HIP(hipMemcpy(sorted_idx, bvh->sorted_idx.p, size_t(n) * 4, hipMemcpyDeviceToHost));
        if (ploc_nodes)
            DPL_HIP(hipMemcpy(ploc_nodes, bvh->pnodes.p, size_t(2 * n - 1) * sizeof(PNode), hipMemcpyDeviceToHost));
        if (iter_counts) std::memcpy(iter_counts, bvh->iter_counts.data(), bvh->iter_counts.size() * 4);
    });
}

static dpl_status TraceImpl(dpl_context* ctx, const dpl_bvh* bvh, const dpl_ray* rays, int64_t n, uint32_t flags,
                            dpl_hit* hits, dpl_mem mem, int64_t* counts_out) {
    if (!ctx || !bvh) return DPL_ERR_INVALID_ARG;
    return Guard(ctx, [&] {
        Require(n >= 0, "n must be non-negative");
        Require(n == 0 || (rays && hits), "rays and hits must be non-null");
        Require((flags & ~(DPL_TRACE_ANY | DPL_TRACE_CULL_BACK | DPL_TRACE_WIDE | DPL_TRACE_COHERENT)) == 0u,
                "unknown trace flags");
        Require(mem == DPL_MEM_HOST || mem == DPL_MEM_DEVICE, "mem must be DPL_MEM_HOST or DPL_MEM_DEVICE");
        Require(bvh->ctx == ctx, "bvh belongs to another context");
        Require(!bvh->batched, "a batched BVH is traced with dpl_trace_batched");
        const bool wide = flags & DPL_TRACE_WIDE;
        Require(!wide || bvh->wide_width != 0, "DPL_TRACE_WIDE needs a wide tree (dpl_bvh_widen)");
        if (counts_out) counts_out[0] = counts_out[1] = 0;
        if (n == 0) return;
        hipStream_t s = ctx->stream;
        const dpl_ray* dr = rays;
        dpl_hit* dh = hits;
        if (mem == DPL_MEM_HOST) {
            dr = ctx->rays.ensure<dpl_ray>(size_t(n) * sizeof(dpl_ray));
            dh = ctx->hits.ensure<dpl_hit>(size_t(n) * sizeof(dpl_hit));
            DPL_HIP(hipMemcpyAsync(const_cast<dpl_ray*>(dr), rays, size_t(n) * sizeof(dpl_ray), hipMemcpyHostToDevice, s));
        }
        uint8_t* tc = ctx->trace_counter.as<uint8_t>();
        unsigned long long* counters = nullptr;
        if (counts_out) counters = reinterpret_cast<unsigned long long*>(tc + 64);  // zeroed by the launch
        // a coherent tracing order on request (the per-ray results do not depend on it; the
        // counter kernels trace in the given order)
        const int32_t* order = nullptr;
        if ((flags & DPL_TRACE_COHERENT) && !counts_out && TraceRefillLanes() > 0) {
            uint32_t* buf = ctx->ray_order.ensure<uint32_t>(size_t(n) * 16);
            const size_t tb = RayOrderTempBytes(n);
            void* tmp = ctx->ray_order_tmp.ensure<uint8_t>(tb);
            DPL_HIP(LaunchRayOrder(dr, n, bvh->root_lo, bvh->root_hi, buf, tmp, tb, &order, s));
        }
        flags &= ~DPL_TRACE_COHERENT;
        hipError_t e;
        if (wide) {
            const WideTree wt{bvh->wnodes.as<float4>(), bvh->tri_ordered.as<float4>(), bvh->node_count,
                              int32_t(bvh->prim_count), bvh->wide_width, bvh->tree_height};
            int32_t* spill = ctx->spill.ensure<int32_t>(
                std::max<size_t>(16, TraceWideSpillBytes(ctx->cu_count, bvh->wide_width, bvh->tree_height)));
            e = LaunchTraceWide(wt, dr, n, flags & ~DPL_TRACE_WIDE, dh, ctx->cu_count, reinterpret_cast<uint32_t*>(tc),
                                spill, counters, s, order);
        } else {
            TraceTree tree{bvh->fnodes.as<FNode>(), bvh->tri_ordered.as<float4>(), bvh->node_count,
                           int32_t(bvh->prim_count), bvh->tree_height, bvh->ndirs, bvh->sobb_dirs.as<float>()};
            int32_t* spill = ctx->spill.ensure<int32_t>(TraceSpillBytes(ctx->cu_count, bvh->tree_height));
            e = LaunchTrace(tree, dr, n, flags, dh, ctx->cu_count, reinterpret_cast<uint32_t*>(tc), spill, counters,
                            s, 0, order);
        }
        if (counts_out) {
            DPL_HIP(hipMemcpyAsync(counts_out, counters, 16, hipMemcpyDeviceToHost, s));
            DPL_HIP(hipStreamSynchronize(s));
        }
        if (e == hipErrorNotSupported) throw Error(DPL_ERR_UNSUPPORTED, "tree height exceeds the traversal stack");
        DPL_HIP(e);
        if (mem == DPL_MEM_HOST) {
            DPL_HIP(hipMemcpyAsync(hits, dh, size_t(n) * sizeof(dpl_hit), hipMemcpyDeviceToHost, s));
            DPL_HIP(hipStreamSynchronize(s));
        }
    });
}

dpl_status dpl_trace(dpl_context* ctx, const dpl_bvh* bvh, const dpl_ray* rays, int64_t n, uint32_t flags,
                     dpl_hit* hits, dpl_mem mem) {
    return TraceImpl(ctx, bvh, rays, n, flags, hits, mem, nullptr);
}

dpl_status dpl_trace_noalloc(dpl_context* ctx, const dpl_bvh* bvh, const dpl_ray* rays, int64_t n, uint32_t flags,
                             dpl_hit* hits) {
    if (!ctx || !bvh) return ctx ? Fail(ctx, DPL_ERR_INVALID_ARG, "null argument") : DPL_ERR_INVALID_ARG;
    return Guard(ctx, [&] {
        Require(n >= 0, "n must be non-negative");
        Require(n == 0 || (rays && hits), "rays and hits must be non-null device pointers");
        Require((flags & ~(DPL_TRACE_ANY | DPL_TRACE_CULL_BACK)) == 0u,
                "dpl_trace_noalloc takes DPL_TRACE_ANY and DPL_TRACE_CULL_BACK only");
        Require(bvh->ctx == ctx, "bvh belongs to another context");
        Require(!bvh->batched, "dpl_trace_noalloc traces single-scene BVHs");
        Require(ctx->noalloc_scratch != nullptr, "no scratch address for this device");
        if (n == 0) return;
        const size_t spill_bytes = kNoAllocScratchBytes - kNoAllocCounterBytes;
        const int grid = TraceGridFor(ctx->cu_count, n, bvh->tree_height, spill_bytes);
        Require(grid >= 1, "the tree is too tall for the no-allocation scratch");
        uint32_t* counter = reinterpret_cast<uint32_t*>(ctx->noalloc_scratch);
        int32_t* spill = reinterpret_cast<int32_t*>(ctx->noalloc_scratch + kNoAllocCounterBytes);
        TraceTree tree{bvh->fnodes.as<FNode>(), bvh->tri_ordered.as<float4>(), bvh->node_count,
                       int32_t(bvh->prim_count), bvh->tree_height, bvh->ndirs, bvh->sobb_dirs.as<float>()};
        DPL_HIP(LaunchTrace(tree, rays, n, flags, hits, ctx->cu_count, counter, spill, nullptr, ctx->stream, grid));
    });
}

dpl_status dpl_trace_counters(dpl_context* ctx, const dpl_bvh* bvh, const dpl_ray* rays, int64_t n, uint32_t flags,
                              dpl_hit* hits, dpl_mem mem, int64_t* node_visits, int64_t* triangle_tests) {
    if (!node_visits || !triangle_tests) return ctx ? Fail(ctx, DPL_ERR_INVALID_ARG, "null counter") : DPL_ERR_INVALID_ARG;
    int64_t c[2] = {0, 0};
    const dpl_status st = TraceImpl(ctx, bvh, rays, n, flags, hits, mem, c);
    *node_visits = c[0];
    *triangle_tests = c[1];
    return st;
}

dpl_status dpl_trace_batched(dpl_context* ctx, const dpl_bvh* bvh, const dpl_ray* rays, const int32_t* ray_mesh,
                             int64_t n, uint32_t flags, dpl_hit* hits, dpl_mem mem) {
    if (!ctx || !bvh) return ctx ? Fail(ctx, DPL_ERR_INVALID_ARG, "null argument") : DPL_ERR_INVALID_ARG;
    return Guard(ctx, [&] {
        Require(n >= 0, "n must be non-negative");
        Require(n == 0 || (rays && hits && ray_mesh), "rays, ray_mesh and hits must be non-null");
        Require((flags & ~(DPL_TRACE_ANY | DPL_TRACE_CULL_BACK | DPL_TRACE_WIDE)) == 0u, "unknown trace flags");
        Require(mem == DPL_MEM_HOST || mem == DPL_MEM_DEVICE, "mem must be DPL_MEM_HOST or DPL_MEM_DEVICE");
        Require(bvh->ctx == ctx, "bvh belongs to another context");
        Require(bvh->batched, "dpl_trace_batched needs a batched BVH");
        const bool wide = flags & DPL_TRACE_WIDE;
        Require(!wide || bvh->wide_width != 0, "DPL_TRACE_WIDE needs a wide tree (dpl_bvh_widen)");
        if (n == 0) return;
        hipStream_t s = ctx->stream;
        const dpl_ray* dr = rays;
        const int32_t* dm = ray_mesh;
        dpl_hit* dh = hits;
        if (mem == DPL_MEM_HOST) {
            dr = ctx->rays.ensure<dpl_ray>(size_t(n) * sizeof(dpl_ray));
            dm = ctx->ray_mesh.ensure<int32_t>(size_t(n) * 4);
            dh = ctx->hits.ensure<dpl_hit>(size_t(n) * sizeof(dpl_hit));
            DPL_HIP(hipMemcpyAsync(const_cast<dpl_ray*>(dr), rays, size_t(n) * sizeof(dpl_ray), hipMemcpyHostToDevice, s));
            DPL_HIP(hipMemcpyAsync(const_cast<int32_t*>(dm), ray_mesh, size_t(n) * 4, hipMemcpyHostToDevice, s));
        }
        if (wide) {
            int32_t* spill = ctx->spill.ensure<int32_t>(
                std::max<size_t>(16, TraceWideSpillBytes(ctx->cu_count, bvh->wide_width, bvh->tree_height)));
            const WideTree wt{bvh->wnodes.as<float4>(), bvh->tri_ordered.as<float4>(), 0, 0, bvh->wide_width,
                              bvh->tree_height};
            DPL_HIP(LaunchTraceWideBatched(wt, bvh->batch_tree.as<BatchTree>(), bvh->mesh_count, dr, dm, n,
                                           flags & ~DPL_TRACE_WIDE, dh, ctx->cu_count,
                                           ctx->trace_counter.as<uint32_t>(), spill, s));
        } else {
            int32_t* spill = ctx->spill.ensure<int32_t>(TraceSpillBytes(ctx->cu_count, bvh->tree_height));
            DPL_HIP(LaunchTraceBatched(bvh->fnodes.as<FNode>(), bvh->tri_ordered.as<float4>(),
                                       bvh->batch_tree.as<BatchTree>(), bvh->mesh_count, bvh->tree_height, dr, dm, n,
                                       flags, dh, ctx->cu_count, ctx->trace_counter.as<uint32_t>(), spill, s));
        }
        if (mem == DPL_MEM_HOST) {
            DPL_HIP(hipMemcpyAsync(hits, dh, size_t(n) * sizeof(dpl_hit), hipMemcpyDeviceToHost, s));
            DPL_HIP(hipStreamSynchronize(s));
        }
    });
}

// ---- trace passes captured in a hipGraph ----
struct dpl_trace_graph {
    dpl_context* ctx = nullptr;
    const dpl_bvh* bvh = nullptr;  // the traced BVH and its generation at capture
    uint64_t generation = 0;
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DevBuf counter, spill;
};

dpl_status dpl_trace_graph_create(dpl_context* ctx, const dpl_bvh* bvh, const dpl_ray* rays, const int32_t* ray_mesh,
                                  int64_t n, uint32_t flags, dpl_hit* hits, dpl_trace_graph** out) {
    if (!ctx || !bvh || !out) return ctx ? Fail(ctx, DPL_ERR_INVALID_ARG, "null argument") : DPL_ERR_INVALID_ARG;
    *out = nullptr;
    std::unique_ptr<dpl_trace_graph> g(new (std::nothrow) dpl_trace_graph());
    if (!g) return Fail(ctx, DPL_ERR_OOM, "host allocation failed");
    g->ctx = ctx;
    dpl_status st = Guard(ctx, [&] {
        Require(n >= 1 && rays && hits, "a captured trace needs n >= 1 and device rays / hits");
        Require((flags & ~(DPL_TRACE_ANY | DPL_TRACE_CULL_BACK | DPL_TRACE_WIDE)) == 0u, "unknown trace flags");
        Require(bvh->ctx == ctx, "bvh belongs to another context");
        Require(bvh->batched == (ray_mesh != nullptr), "ray_mesh is required exactly for a batched BVH");
        const bool wide = flags & DPL_TRACE_WIDE;
        Require(!wide || bvh->wide_width != 0, "DPL_TRACE_WIDE needs a wide tree (dpl_bvh_widen)");
        const auto live = ctx->bvh_gen.find(bvh);
        Require(live != ctx->bvh_gen.end(), "the BVH has no built tree");
        g->bvh = bvh;
        g->generation = live->second;
        uint32_t* counter = g->counter.ensure<uint32_t>(kTraceCounterBytes);
        const size_t spill_bytes = wide ? TraceWideSpillBytes(ctx->cu_count, bvh->wide_width, bvh->tree_height)
                                        : TraceSpillBytes(ctx->cu_count, bvh->tree_height);
        int32_t* spill = g->spill.ensure<int32_t>(std::max<size_t>(16, spill_bytes));
        DPL_HIP(hipStreamCreateWithFlags(&g->capture, hipStreamNonBlocking));
        DPL_HIP(hipStreamBeginCapture(g->capture, hipStreamCaptureModeThreadLocal));
        hipError_t e;
        if (bvh->batched && wide) {
            const WideTree wt{bvh->wnodes.as<float4>(), bvh->tri_ordered.as<float4>(), 0, 0, bvh->wide_width,
                              bvh->tree_height};
            e = LaunchTraceWideBatched(wt, bvh->batch_tree.as<BatchTree>(), bvh->mesh_count, rays, ray_mesh, n,
                                       flags & ~DPL_TRACE_WIDE, hits, ctx->cu_count, counter, spill, g->capture);
        } else if (bvh->batched) {
            e = LaunchTraceBatched(bvh->fnodes.as<FNode>(), bvh->tri_ordered.as<float4>(),
                                   bvh->batch_tree.as<BatchTree>(), bvh->mesh_count, bvh->tree_height, rays, ray_mesh,
                                   n, flags, hits, ctx->cu_count, counter, spill, g->capture);
        } else if (wide) {
            const WideTree wt{bvh->wnodes.as<float4>(), bvh->tri_ordered.as<float4>(), bvh->node_count,
                              int32_t(bvh->prim_count), bvh->wide_width, bvh->tree_height};
            e = LaunchTraceWide(wt, rays, n, flags & ~DPL_TRACE_WIDE, hits, ctx->cu_count, counter, spill, nullptr,
                                g->capture);
        } else {
            TraceTree tree{bvh->fnodes.as<FNode>(), bvh->tri_ordered.as<float4>(), bvh->node_count,
                           int32_t(bvh->prim_count), bvh->tree_height, bvh->ndirs, bvh->sobb_dirs.as<float>()};
            e = LaunchTrace(tree, rays, n, flags, hits, ctx->cu_count, counter, spill, nullptr, g->capture);
        }
        hipGraph_t graph = nullptr;
        const hipError_t e2 = hipStreamEndCapture(g->capture, &graph);
        g->graph = graph;
        DPL_HIP(e);
        DPL_HIP(e2);
        DPL_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    });
    if (st != DPL_OK) {
        dpl_trace_graph_destroy(g.release());
        return st;
    }
    ctx->graphs.insert(g.get());
    *out = g.release();
    return DPL_OK;
}

dpl_status dpl_trace_graph_launch(dpl_trace_graph* g) {
    if (!g || !g->exec || !g->ctx) return DPL_ERR_INVALID_ARG;
    return Guard(g->ctx, [&] {
        // the graph holds the BVH's device addresses, node counts and a spill buffer sized
        // for its height at capture: replay only the tree it was captured from
        const auto live = g->ctx->bvh_gen.find(g->bvh);
        Require(live != g->ctx->bvh_gen.end() && live->second == g->generation,
                "the BVH was rebuilt, widened or destroyed after the graph was captured");
        DPL_HIP(hipGraphLaunch(g->exec, g->ctx->stream));
    });
}

static void DetachGraphs(dpl_context* ctx) {
    for (dpl_trace_graph* g : ctx->graphs) g->ctx = nullptr;
    ctx->graphs.clear();
}

void dpl_trace_graph_destroy(dpl_trace_graph* g) {
    if (!g) return;
    if (g->ctx) {
        (void)hipSetDevice(g->ctx->device);
        (void)hipStreamSynchronize(g->ctx->stream);
        g->ctx->graphs.erase(g);
    }
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->capture) (void)hipStreamDestroy(g->capture);
    delete g;
}

// ---- wide trees ----
static int64_t NodeSlots(const dpl_bvh* bvh) { return bvh->batched ? bvh->fnode_slots : bvh->node_count; }

static void WideInfo(const dpl_bvh* bvh, dpl_wide_info* info) {
    info->width = bvh->wide_width;
    info->node_bytes = bvh->wide_width ? WideNodeBytes(bvh->wide_width) : 0;
    info->node_count = bvh->wide_width ? NodeSlots(bvh) : 0;
    info->widen_ms = bvh->wide_width ? bvh->widen_ms : 0.0;
}

dpl_status dpl_bvh_widen(dpl_bvh* bvh, int32_t width, dpl_wide_info* info) {
    if (!bvh || !bvh->ctx) return DPL_ERR_INVALID_ARG;
    dpl_context* ctx = bvh->ctx;
    return Guard(ctx, [&] {
        Require(width == 4 || width == 8, "width must be 4 or 8");
        Require(NodeSlots(bvh) <= INT32_MAX, "too many nodes");
        BumpGeneration(bvh);
        hipStream_t s = ctx->stream;
        const int32_t K = int32_t(NodeSlots(bvh));
        bvh->wide_width = 0;
        float4* out = bvh->wnodes.ensure<float4>(std::max<size_t>(16, size_t(K) * WideNodeBytes(width)));
        const FNode* src = bvh->fnodes.as<FNode>();
        FNode* refit = nullptr;
        float* box = nullptr;
        int32_t* done = nullptr;
        if (bvh->ndirs > 0) {  // SOBB child slabs: refit axis-aligned boxes first
            uint8_t* t = ctx->wide_tmp.ensure<uint8_t>(std::max<size_t>(16, size_t(K) * (sizeof(FNode) + 28)));
            refit = reinterpret_cast<FNode*>(t);
            box = reinterpret_cast<float*>(t + size_t(K) * sizeof(FNode));
            done = reinterpret_cast<int32_t*>(box + 6 * size_t(K));
        }
        DPL_HIP(hipEventRecord(ctx->ev[0], s));
        if (refit) {
            DPL_HIP(LaunchRefitAabb(src, K, bvh->tri_ordered.as<float4>(), int32_t(bvh->prim_count), bvh->tree_height,
                                    refit, box, done, s));
            src = refit;
        }
        DPL_HIP(LaunchWiden(src, K, width, out, bvh->batched ? bvh->batch_tree.as<BatchTree>() : nullptr,
                            bvh->batched ? bvh->mesh_count : 0, s));
        DPL_HIP(hipEventRecord(ctx->ev[1], s));
        DPL_HIP(hipEventSynchronize(ctx->ev[1]));
        float ms = 0.0f;
        DPL_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        bvh->widen_ms = ms;
        bvh->wide_width = width;
        if (info) WideInfo(bvh, info);
    });
}

dpl_status dpl_bvh_get_wide(const dpl_bvh* bvh, dpl_wide_info* info) {
    if (!bvh || !info) return DPL_ERR_INVALID_ARG;
    WideInfo(bvh, info);
    return DPL_OK;
}

dpl_status dpl_bvh_download_wide(const dpl_bvh* bvh, void* nodes) {
    if (!bvh || !bvh->ctx || !nodes) return DPL_ERR_INVALID_ARG;
    dpl_context* ctx = bvh->ctx;
    return Guard(ctx, [&] {
        Require(bvh->wide_width != 0, "the BVH has no wide tree (dpl_bvh_widen)");
        DPL_HIP(hipStreamSynchronize(ctx->stream));
        DPL_HIP(hipMemcpy(nodes, bvh->wnodes.p, size_t(NodeSlots(bvh)) * WideNodeBytes(bvh->wide_width),
                          hipMemcpyDeviceToHost));
    });
}

dpl_status dpl_compute_codes(dpl_context* ctx, const float* tri, int64_t n, int32_t morton_type, uint64_t* codes,
                             float* bounds12) {
    if (!ctx) return DPL_ERR_INVALID_ARG;
    return Guard(ctx, [&] {
        Require(tri && codes && n >= 1, "invalid arguments");
        Require(morton_type == 0 || morton_type == 1, "morton_type must be 0 or 1");
        hipStream_t s = ctx->stream;
        BuildBuffers b{};
        b.n = n;
        b.tri = tri;
        b.keys = codes;
        b.vals = ctx->vals.ensure<uint32_t>(size_t(n) * 4);
        uint8_t* sm = ctx->small.ensure<uint8_t>(1024 + size_t(kBoundsBlocks) * 64);
        b.bounds = reinterpret_cast<uint32_t*>(sm);
        b.bounds_part = reinterpret_cast<uint32_t*>(sm + 1024);
        // the init kernel touches these; give it valid scratch
        b.tile_counter = reinterpret_cast<uint32_t*>(sm + 64);
        b.iter_done = reinterpret_cast<int32_t*>(sm + 96);
        b.counters = reinterpret_cast<int32_t*>(sm + 128);
        b.sah = reinterpret_cast<double*>(sm + 192);
        b.ccount = ctx->ccount.ensure<int32_t>(size_t(n + 2) * 4);
        b.parent = ctx->parent.ensure<int32_t>(size_t(2 * n - 1) * 4);
        b.status[0] = ctx->status0.ensure<uint64_t>(size_t(PlocStatusWords(n, 256)) * 8);
        b.leafslot = ctx->leafslot.ensure<int32_t>(size_t(n) * 4);
        DPL_HIP(LaunchBuildInit(b, 256, s));
        DPL_HIP(LaunchBounds(b, s));
        DPL_HIP(LaunchCodes(b, morton_type, s));
        if (bounds12) {
            uint32_t raw[13];
            DPL_HIP(hipMemcpyAsync(raw, b.bounds, sizeof(raw), hipMemcpyDeviceToHost, s));
            DPL_HIP(hipStreamSynchronize(s));
            for (int k = 0; k < 12; k++) {
                uint32_t key = raw[k];
                uint32_t u = (key & 0x80000000u) ? (key & 0x7FFFFFFFu) : ~key;
                std::memcpy(&bounds12[k], &u, 4);
            }
        }
    });
}

dpl_status dpl_radix_sort_pairs(dpl_context* ctx, uint64_t* keys, uint32_t* values, int64_t n, int32_t key_bits,
                                int32_t digit_bits) {
    if (!ctx) return DPL_ERR_INVALID_ARG;
    return Guard(ctx, [&] {
        Require(keys && values && n >= 0, "invalid arguments");
        Require(key_bits >= 1 && key_bits <= 64, "key_bits must be in [1, 64]");
        Require(digit_bits == 4 || digit_bits == 8, "digit_bits must be 4 or 8");
        if (n == 0) return;
        const bool key64 = key_bits > 32;
        hipStream_t s = ctx->stream;
        uint64_t* ka = ctx->keys_alt.ensure<uint64_t>(size_t(n) * 8);
        uint32_t* va = ctx->vals_alt.ensure<uint32_t>(size_t(n) * 4);
        const size_t tb = RadixSortTempBytes(n, key_bits, digit_bits, key64);
        void* tmp = ctx->sort_tmp.ensure(tb);
        bool in_alt = false;
        DPL_HIP(LaunchRadixSort(keys, values, ka, va, n, key_bits, digit_bits, key64, tmp, tb, &in_alt, s));
        if (in_alt) {
            DPL_HIP(hipMemcpyAsync(keys, ka, size_t(n) * (key64 ? 8 : 4), hipMemcpyDeviceToDevice, s));
            DPL_HIP(hipMemcpyAsync(values, va, size_t(n) * 4, hipMemcpyDeviceToDevice, s));
        }
    });
}

static dpl_status Render(bool adjoint, dpl_context* ctx, const dpl_bvh* bvh, const dpl_render_params* rp,
                         const int32_t* mat_id, const float* albedo, const float* emission, const float* d_image,
                         float* image, float* grad_albedo, float* grad_emission) {
    if (!ctx || !bvh || !rp) return ctx ? Fail(ctx, DPL_ERR_INVALID_ARG, "null argument") : DPL_ERR_INVALID_ARG;
    return Guard(ctx, [&] {
        Require(bvh->ctx == ctx, "bvh belongs to another context");
        Require(!bvh->batched, "path tracing needs a single-scene BVH");
        Require(rp->width >= 2 && rp->height >= 2, "width and height must be at least 2");
        Require(rp->spp >= 1, "spp must be at least 1");
        Require(rp->max_depth >= 1 && rp->max_depth <= (adjoint ? 16 : 64), "max_depth out of range");
        Require(rp->material_count >= 1 && rp->material_count <= 256, "material_count must be in [1, 256]");
        Require(int64_t(rp->width) * rp->height * rp->spp < (int64_t(1) << 40), "too many paths");
        Require(mat_id && albedo && emission, "material arrays must be non-null");
        if (adjoint) Require(d_image && grad_albedo && grad_emission, "adjoint buffers must be non-null");
        else Require(image != nullptr, "image must be non-null");
        float basis[12];
        CameraBasisHost(rp->camera, rp->width, rp->height, basis);
        TraceTree tree{bvh->fnodes.as<FNode>(), bvh->tri_ordered.as<float4>(), bvh->node_count,
                       int32_t(bvh->prim_count), bvh->tree_height, bvh->ndirs, bvh->sobb_dirs.as<float>()};
        int32_t* spill = ctx->spill.ensure<int32_t>(TraceSpillBytes(ctx->cu_count, bvh->tree_height));
        hipError_t e = LaunchRender(adjoint, tree, *rp, basis, mat_id, albedo, emission, d_image, image, grad_albedo,
                                    grad_emission, ctx->cu_count, ctx->trace_counter.as<uint32_t>(), spill,
                                    ctx->stream);
        if (e == hipErrorNotSupported) throw Error(DPL_ERR_UNSUPPORTED, "tree height exceeds the traversal stack");
        DPL_HIP(e);
    });
}

namespace {

dpl_status WhittedImpl(dpl_context* ctx, const dpl_bvh* bvh, const dpl_whitted_params* wp, const float* tri,
                       const float* vn, const int32_t* mat_id, const dpl_material* mats, int32_t material_count,
                       const int32_t* lights, int32_t light_count, float* image, dpl_whitted_stats* stats) {
    if (!ctx || !bvh || !wp) return ctx ? Fail(ctx, DPL_ERR_INVALID_ARG, "null argument") : DPL_ERR_INVALID_ARG;
    return Guard(ctx, [&] {
        Require(bvh->ctx == ctx, "bvh belongs to another context");
        Require(!bvh->batched, "Whitted rendering needs a single-scene BVH");
        Require(wp->width >= 2 && wp->height >= 2, "width and height must be at least 2");
        Require(wp->max_depth >= 0 && wp->max_depth <= kDplWhittedMaxDepth, "max_depth must be in [0, 8]");
        Require(wp->area_light_sample_cnt >= 0, "area_light_sample_cnt must be non-negative");
        Require(tri && mat_id && mats && image, "triangle, material and image arrays must be non-null");
        Require(material_count >= 1, "at least one material is required");
        Require(light_count >= 0 && (light_count == 0 || lights), "light_tris must be non-null when light_count > 0");
        float basis[12];
        CameraBasisHost(wp->camera, wp->width, wp->height, basis);
        TraceTree tree{bvh->fnodes.as<FNode>(), bvh->tri_ordered.as<float4>(), bvh->node_count,
                       int32_t(bvh->prim_count), bvh->tree_height, bvh->ndirs, bvh->sobb_dirs.as<float>()};
        int32_t* spill = ctx->spill.ensure<int32_t>(std::max<size_t>(16, TraceSpillBytes(ctx->cu_count, bvh->tree_height)));
        uint8_t* tc = ctx->trace_counter.ensure<uint8_t>(kTraceCounterBytes);
        // the six totals sit past the pixel counter; zeroed by the launch
        unsigned long long* totals = stats ? reinterpret_cast<unsigned long long*>(tc + 64) : nullptr;
        hipEvent_t ev[2] = {nullptr, nullptr};
        if (stats) {
            DPL_HIP(hipEventCreate(&ev[0]));
            DPL_HIP(hipEventCreate(&ev[1]));
            DPL_HIP(hipEventRecord(ev[0], ctx->stream));
        }
        hipError_t e = LaunchWhitted(tree, *wp, basis, tri, vn, mat_id, mats, lights, light_count, image, ctx->cu_count,
                                     reinterpret_cast<uint32_t*>(tc), spill, totals, ctx->stream);
        if (stats) {
            unsigned long long v[6] = {};
            hipError_t e2 = hipEventRecord(ev[1], ctx->stream);
            if (e2 == hipSuccess) e2 = hipMemcpyAsync(v, totals, sizeof(v), hipMemcpyDeviceToHost, ctx->stream);
            if (e2 == hipSuccess) e2 = hipStreamSynchronize(ctx->stream);
            float ms = 0.0f;
            if (e2 == hipSuccess) e2 = hipEventElapsedTime(&ms, ev[0], ev[1]);
            (void)hipEventDestroy(ev[0]);
            (void)hipEventDestroy(ev[1]);
            if (e == hipSuccess) DPL_HIP(e2);
            *stats = dpl_whitted_stats{int64_t(v[0]), int64_t(v[1]), int64_t(v[2]), int64_t(v[3]), int64_t(v[4]),
                                       int64_t(v[5]), double(ms)};
        }
        if (e == hipErrorNotSupported) throw Error(DPL_ERR_UNSUPPORTED, "tree height exceeds the traversal stack");
        DPL_HIP(e);
    });
}

}  // namespace

dpl_status dpl_render_whitted(dpl_context* ctx, const dpl_bvh* bvh, const dpl_whitted_params* wp, const float* tri,
                              const float* vn, const int32_t* mat_id, const dpl_material* mats, int32_t material_count,
                              const int32_t* lights, int32_t light_count, float* image) {
    return WhittedImpl(ctx, bvh, wp, tri, vn, mat_id, mats, material_count, lights, light_count, image, nullptr);
}

dpl_status dpl_render_whitted_stats(dpl_context* ctx, const dpl_bvh* bvh, const dpl_whitted_params* wp,
                                    const float* tri, const float* vn, const int32_t* mat_id, const dpl_material* mats,
                                    int32_t material_count, const int32_t* lights, int32_t light_count, float* image,
                                    dpl_whitted_stats* stats) {
    if (!stats) return ctx ? Fail(ctx, DPL_ERR_INVALID_ARG, "null stats") : DPL_ERR_INVALID_ARG;
    return WhittedImpl(ctx, bvh, wp, tri, vn, mat_id, mats, material_count, lights, light_count, image, stats);
}

dpl_status dpl_render_forward(dpl_context* ctx, const dpl_bvh* bvh, const dpl_render_params* rp, const int32_t* mat_id,
                              const float* albedo, const float* emission, float* image) {
    return Render(false, ctx, bvh, rp, mat_id, albedo, emission, nullptr, image, nullptr, nullptr);
}

dpl_status dpl_render_adjoint(dpl_context* ctx, const dpl_bvh* bvh, const dpl_render_params* rp,
                              const int32_t* mat_id, const float* albedo, const float* emission,
                              const float* d_image, float* grad_albedo, float* grad_emission) {
    return Render(true, ctx, bvh, rp, mat_id, albedo, emission, d_image, nullptr, grad_albedo, grad_emission);
}

}  // extern "C"

// ---- autotuner ----
namespace {

struct TuneEntry {
    int radius = 0, block = 0, digit = 0;
    double build_ms = 0.0, sah = 0.0;
};

// The cache is a JSON object written one entry per line by WriteTuneCache:
// {"version": 1, "entries": {
//   "<key>": {"nn_search_radius": 16, "block_size": 256, "sort_digit_bits": 8, "build_ms": 4.5, "sah_cost": 284.5},
// ...}}
std::map<std::string, TuneEntry> ReadTuneCache(const char* path) {
    std::map<std::string, TuneEntry> m;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        const size_t q0 = line.find('"');
        const size_t q1 = q0 == std::string::npos ? q0 : line.find('"', q0 + 1);
        const size_t br = q1 == std::string::npos ? q1 : line.find('{', q1);
        if (br == std::string::npos) continue;
        const std::string key = line.substr(q0 + 1, q1 - q0 - 1);
        if (key == "version" || key == "entries") continue;
        TuneEntry e;
        if (std::sscanf(line.c_str() + br,
                        "{\"nn_search_radius\": %d, \"block_size\": %d, \"sort_digit_bits\": %d, \"build_ms\": %lf, "
                        "\"sah_cost\": %lf}",
                        &e.radius, &e.block, &e.digit, &e.build_ms, &e.sah) == 5)
            m[key] = e;
    }
    return m;
}

void WriteTuneCache(const char* path, const std::map<std::string, TuneEntry>& m) {
    const std::string tmp = std::string(path) + ".tmp";
    {
        std::ofstream f(tmp, std::ios::trunc);
        if (!f) throw Error(DPL_ERR_INVALID_ARG, std::string("cannot write the autotune cache ") + path);
        f << "{\"version\": 1, \"entries\": {\n";
        size_t k = 0;
        for (const auto& [key, e] : m) {
            char buf[512];
            std::snprintf(buf, sizeof(buf),
                          "  \"%s\": {\"nn_search_radius\": %d, \"block_size\": %d, \"sort_digit_bits\": %d, "
                          "\"build_ms\": %.6f, \"sah_cost\": %.9g}%s\n",
                          key.c_str(), e.radius, e.block, e.digit, e.build_ms, e.sah, ++k < m.size() ? "," : "");
            f << buf;
        }
        f << "}}\n";
    }
    if (std::rename(tmp.c_str(), path) != 0) throw Error(DPL_ERR_INVALID_ARG, "cannot replace the autotune cache");
}

int FloorLog2(int64_t v) {
    int r = 0;
    while (v > 1) {
        v >>= 1;
        r++;
    }
    return r;
}

}  // namespace

dpl_status dpl_autotune(dpl_context* ctx, const float* tri, int64_t n, const int64_t* offsets, int32_t mesh_count,
                        dpl_mem mem, const dpl_build_params* base, double sah_tolerance, const char* cache_path,
                        dpl_build_params* best, dpl_autotune_info* info) {
    if (!ctx || !base || !best) return ctx ? Fail(ctx, DPL_ERR_INVALID_ARG, "null argument") : DPL_ERR_INVALID_ARG;
    return Guard(ctx, [&] {
        const dpl_build_params b0 = Normalize(base);
        const bool batched = offsets != nullptr;
        Require(tri != nullptr, "tri_xyz is null");
        Require(sah_tolerance >= 0.0, "sah_tolerance must be non-negative");
        const int64_t N = batched ? (mesh_count >= 1 ? offsets[mesh_count] : 0) : n;
        Require(N >= 1, "at least one triangle is required");
        std::ostringstream key;
        key << (batched ? "batched" : "single") << "/n2^" << FloorLog2(N) << "/m2^"
            << FloorLog2(batched ? mesh_count : 1) << "/" << (b0.morton_type == DPL_EMC60 ? "emc60" : "morton30") << "/"
            << (b0.bv_type == DPL_BV_SOBB ? "sobb" + std::to_string(b0.kdop_size) : std::string("aabb")) << "/leaf"
            << b0.max_triangles_per_leaf << "/" << ctx->device_name;
        std::map<std::string, TuneEntry> cache;
        if (cache_path) cache = ReadTuneCache(cache_path);
        dpl_autotune_info inf{};
        auto hit = cache.find(key.str());
        if (hit != cache.end()) {
            *best = b0;
            best->nn_search_radius = hit->second.radius;
            best->block_size = hit->second.block;
            best->sort_digit_bits = hit->second.digit;
            inf.build_ms = hit->second.build_ms;
            inf.sah_cost = hit->second.sah;
            inf.from_cache = 1;
            if (info) *info = inf;
            return;
        }
        // stage host input once so that the timed builds read device memory
        const float* dtri = tri;
        DevBuf staged;
        if (mem == DPL_MEM_HOST) {
            float* d = staged.ensure<float>(size_t(N) * 36);
            DPL_HIP(hipMemcpy(d, tri, size_t(N) * 36, hipMemcpyHostToDevice));
            dtri = d;
        }
        dpl_bvh scratch;
        scratch.ctx = ctx;
        std::vector<std::pair<dpl_build_params, TuneEntry>> runs;
        const std::vector<int> blocks = batched ? std::vector<int>{256, 512} : std::vector<int>{128, 256, 512};
        for (int r : {8, 16, 32})
            for (int blk : blocks)
                for (int dg : {4, 8}) {
                    dpl_build_params q = b0;
                    q.nn_search_radius = r;
                    q.block_size = blk;
                    q.sort_digit_bits = dg;
                    q.keep_intermediates = 0;
                    double ms = 1e30, sah = 0.0;
                    for (int rep = 0; rep < 3; rep++) {  // the first build warms up
                        dpl_build_stats st{};
                        if (batched) BatchImpl(ctx, &scratch, dtri, offsets, mesh_count, DPL_MEM_DEVICE, q, &st);
                        else BuildImpl(ctx, &scratch, dtri, n, DPL_MEM_DEVICE, q, &st);
                        if (rep > 0) ms = std::min(ms, st.build_ms);
                        sah = st.sah_cost;
                    }
                    runs.push_back({q, TuneEntry{r, blk, dg, ms, sah}});
                }
        double min_sah = 1e300;
        for (const auto& r : runs) min_sah = std::min(min_sah, r.second.sah);
        const double limit = min_sah * (1.0 + sah_tolerance);
        size_t pick = runs.size();
        for (size_t k = 0; k < runs.size(); k++)
            if (runs[k].second.sah <= limit && (pick == runs.size() || runs[k].second.build_ms < runs[pick].second.build_ms))
                pick = k;
        *best = runs[pick].first;
        best->keep_intermediates = b0.keep_intermediates;
        inf.build_ms = runs[pick].second.build_ms;
        inf.sah_cost = runs[pick].second.sah;
        inf.candidates = int32_t(runs.size());
        if (cache_path) {
            cache[key.str()] = runs[pick].second;
            WriteTuneCache(cache_path, cache);
        }
        if (info) *info = inf;
    });
}
