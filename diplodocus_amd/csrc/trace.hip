// Ray-batch trace kernel for gfx950: persistent grid, rays fetched 64 at a time per wave
// from a global atomic counter, an LDS stack ring per lane (16 entries; 8 for SOBB trees in
// the wave-at-a-time kernel) with a global spill region for deeper stacks (see traverse.h),
// AABB or SOBB child boxes, optional traversal counters. The default kernel replaces a
// lane's finished ray at once and keeps up to four leaves pending per lane.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "dpl_internal.h"
#include "traverse.h"

namespace dpl {

namespace {
constexpr int kStackLds = 16;
// SOBB kernels keep 1/(n.d) of all 13 directions per lane in LDS (52 B, n.o is recomputed).
// For the wave-at-a-time kernel a shorter LDS stack ring keeps more workgroups resident per
// CU, which pays more than the extra spills. The per-lane refill kernel keeps 16 entries (2M
// mixed scene, 4.2M random rays, SOBB tree: 4 entries 25.5 ms, 8 21.6 ms, 16 16.0 ms, 32
// 19.0 ms; 8 entries spill ~2.6 KB per ray through L2 to the fabric). Keeping n.o per
// direction in LDS as well gives the same time with 8 entries (fewer VALU instructions, fewer
// resident waves) and 17.9 ms with 16. AABB trees: 8 entries 30.1 ms, 16 20.5, 32 20.2.
constexpr int kStackLdsSobb = 8;
constexpr int kStackLdsSobbRefill = 16;
constexpr bool kSobbSpillInterleaved = true;
constexpr int kTraceBlock = 256;
constexpr int kTraceWgPerCu = 8;
constexpr int kTraceRefillDefault = 8;
// leaves a lane keeps pending while its wave is among inner nodes: 2M-triangle random rays
// 26.9 ms with 1, 22.4 with 2, 21.5 with 4, 21.6 with 8
constexpr int kTracePending = 4;  // 2M-triangle random rays: 8 -> 26.3 ms, 16 -> 26.7, 32 -> 28.5, wave-at-a-time 28.1
}  // namespace

// BATCH: ray i is traced against mesh ray_mesh[i], whose tree starts at the per-mesh offsets
// of mt (tree.nodes / tree.tris are then the bases of all meshes' nodes and triangles).
template <int BLOCK, bool COUNT, bool SOBB, bool BATCH>
__global__ __launch_bounds__(BLOCK) void KTrace(TraceTree tree, const float4* __restrict__ rays, int64_t n,
                                                uint32_t flags, dpl_hit* __restrict__ hits, uint32_t* counter,
                                                int32_t* spill, int32_t spill_stride,
                                                unsigned long long* counters, const int32_t* __restrict__ ray_mesh,
                                                const BatchTree* __restrict__ mt, int32_t mesh_count) {
    constexpr int S = SOBB ? kStackLdsSobb : kStackLds;
    __shared__ int32_t s_stack[S][BLOCK];
    __shared__ uint32_t s_base[BLOCK / kWave];
    __shared__ float s_idn[SOBB ? kMaxDirs : 1][BLOCK];
    __shared__ float s_dirs[SOBB ? 3 * kMaxDirs : 1];
    if (SOBB) {
        for (int k = threadIdx.x; k < 3 * tree.ndirs; k += BLOCK) s_dirs[k] = tree.dirs[k];
        __syncthreads();
    }
    const bool any = flags & DPL_TRACE_ANY, cull = flags & DPL_TRACE_CULL_BACK;
    const int lane = LaneId(), wid = threadIdx.x / kWave;
    SobbRay sr{s_dirs, &s_idn[0][threadIdx.x], 0.0f, 0.0f, 0.0f, 0.0f};
    TraceCounters cnt{0u, 0u};
    auto run = [&](auto& st) {
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
                if (SOBB) PrepareSobb<BLOCK>(sr, s_dirs, tree.ndirs, r0.x, r0.y, r0.z, r1.x, r1.y, r1.z);
                TraceTree t = tree;
                bool valid = true;
                if (BATCH) {
                    const int32_t m = ray_mesh[i];
                    valid = m >= 0 && m < mesh_count;
                    if (valid) {
                        const BatchTree bt = mt[m];
                        t.nodes = tree.nodes + bt.node_off;
                        t.tris = tree.tris + 3 * bt.tri_off;
                        t.node_count = bt.node_count;
                        t.prim_count = bt.prim_count;
                        t.height = bt.height;
                    }
                }
                HitResult h{0.0f, 0.0f, 0.0f, -1, -1};
                if (valid)
                    h = any ? Traverse<true, SOBB>(t, w, sr, cull, st, cnt) : Traverse<false, SOBB>(t, w, sr, cull, st, cnt);
                dpl_hit o;
                o.t = h.t;
                o.b1 = h.b1;
                o.b2 = h.b2;
                o.prim = h.prim;
                hits[i] = o;
            }
            __builtin_amdgcn_wave_barrier();
        }
    };
    if constexpr (SOBB && kSobbSpillInterleaved) {
        LaneStackInterleaved<S, BLOCK> st{&s_stack[0][threadIdx.x], spill + int64_t(blockIdx.x) * BLOCK + threadIdx.x,
                                          int64_t(gridDim.x) * BLOCK, 0};
        run(st);
    } else {
        LaneStack<S, BLOCK> st{&s_stack[0][threadIdx.x],
                               spill + (int64_t(blockIdx.x) * BLOCK + threadIdx.x) * spill_stride, 0};
        run(st);
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

// The same trace with each lane replacing its ray as soon as it terminates and leaves
// postponed while the wave is still among inner nodes (persistent threads, "while-while" with
// speculative traversal and replacement of terminated rays: Aila & Laine, HPG 2009). The
// wave fetches rays for all its idle lanes with one atomic once at least `refill` of them are
// idle. Every lane runs the node loop while any lane is still looking for its first leaf; a
// lane that meets a leaf keeps it pending (up to Q of them) and goes on with the next stack
// entry, and then the pending leaves are tested one triangle per step, so that a lane with
// several small leaves is not held to the slowest lane of each leaf. Closest hits are the
// smallest t in (tmin, tmax), ties to the lower primitive id, and culling by the best t so far
// is conservative, so the result does not depend on the visiting order: it equals the ordered
// traversal's bit for bit.
template <int BLOCK, bool SOBB, bool BATCH, bool ANY, int Q>
__global__ __launch_bounds__(BLOCK) void KTraceRefill(TraceTree tree, const float4* __restrict__ rays, int64_t n,
                                                      bool cull, dpl_hit* __restrict__ hits, uint32_t* counter,
                                                      int32_t* spill, int32_t spill_stride,
                                                      const int32_t* __restrict__ ray_mesh,
                                                      const BatchTree* __restrict__ mt, int32_t mesh_count,
                                                      int refill, const int32_t* __restrict__ order,
                                                      int shares, int64_t share_len) {
    constexpr int S = SOBB ? kStackLdsSobbRefill : kStackLds;
    constexpr int32_t kEnd = 0x7FFFFFFF;  // no next node: the ray is done once no leaf is pending
    __shared__ int32_t s_stack[S][BLOCK];
    __shared__ uint32_t s_base[BLOCK / kWave];
    __shared__ float s_idn[SOBB ? kMaxDirs : 1][BLOCK];
    __shared__ float s_dirs[SOBB ? 3 * kMaxDirs : 1];
    if (SOBB) {
        for (int k = threadIdx.x; k < 3 * tree.ndirs; k += BLOCK) s_dirs[k] = tree.dirs[k];
        __syncthreads();
    }
    const int lane = LaneId(), wid = threadIdx.x / kWave;
    const uint64_t below = (1ull << lane) - 1ull;
    SobbRay sr{s_dirs, &s_idn[0][threadIdx.x], 0.0f, 0.0f, 0.0f, 0.0f};
    auto run = [&](auto& st) {
        int64_t ray = -1;  // this lane's ray, -1 while idle
        WtRay w{};
        TraceTree t = tree;
        float best_t = 0.0f;
        HitResult r{0.0f, 0.0f, 0.0f, -1, -1};
        int32_t ref = kEnd;
        int32_t pq[Q];  // leaves waiting for their test (pq[0] newest, pq[npend - 1] oldest)
        int npend = 0;
#pragma unroll
        for (int q = 0; q < Q; q++) pq[q] = 0;
        int64_t budget = 0;
        bool more = true;  // wave-uniform: the ray counter may still hand out rays
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
                        if (SOBB) PrepareSobb<BLOCK>(sr, s_dirs, tree.ndirs, r0.x, r0.y, r0.z, r1.x, r1.y, r1.z);
                        t = tree;
                        ref = 0;
                        if (BATCH) {
                            const int32_t m = ray_mesh[i];
                            if (m >= 0 && m < mesh_count) {
                                const BatchTree bt = mt[m];
                                t.nodes = tree.nodes + bt.node_off;
                                t.tris = tree.tris + 3 * bt.tri_off;
                                t.node_count = bt.node_count;
                                t.prim_count = bt.prim_count;
                                t.height = bt.height;
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
                // inner nodes, while any lane has no pending leaf yet (a lane keeps up to Q)
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
                        const float4* np = reinterpret_cast<const float4*>(t.nodes + ref);
                        const float4 a = np[0], b = np[1], c = np[2];
                        const int4 d = reinterpret_cast<const int4*>(np)[3];
                        float tl, tr;
                        bool hl, hr;
                        if (SOBB) {
                            const uint32_t dirs = uint32_t(d.z);
                            hl = SobbBox<BLOCK>(sr, w.tmin, best_t, a.x, a.y, a.z, a.w, b.x, b.y, dirs & 0xFFFu, tl);
                            hr = SobbBox<BLOCK>(sr, w.tmin, best_t, b.z, b.w, c.x, c.y, c.z, c.w,
                                                (dirs >> 12) & 0xFFFu, tr);
                        } else {
                            hl = Box(w, best_t, a.x, a.y, a.z, a.w, c.x, c.y, tl);
                            hr = Box(w, best_t, b.x, b.y, b.z, b.w, c.z, c.w, tr);
                        }
                        if (hl && hr) {
                            const bool rfirst = tr < tl;
                            st.Push(rfirst ? d.x : d.y);
                            ref = rfirst ? d.y : d.x;
                        } else if (hl) {
                            ref = d.x;
                        } else if (hr) {
                            ref = d.y;
                        } else {
                            ref = st.sp > 0 ? st.Pop() : kEnd;
                        }
                    }
                }
                // the pending leaves
                bool done = false;
                // one triangle per step over all of the lane's pending leaves, so that a lane
                // with several small leaves is not held to the slowest lane per leaf
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
    };
    if constexpr (SOBB && kSobbSpillInterleaved) {
        LaneStackInterleaved<S, BLOCK> st{&s_stack[0][threadIdx.x], spill + int64_t(blockIdx.x) * BLOCK + threadIdx.x,
                                          int64_t(gridDim.x) * BLOCK, 0};
        run(st);
    } else {
        LaneStack<S, BLOCK> st{&s_stack[0][threadIdx.x],
                               spill + (int64_t(blockIdx.x) * BLOCK + threadIdx.x) * spill_stride, 0};
        run(st);
    }
}

// Idle lanes per wave before the wave fetches new rays (1..64); 0 selects the wave-at-a-time
// kernel (each wave fetches 64 rays and fetches again once all 64 are done).
int TraceRefillLanes() {
    static const int v = [] {
        const char* e = std::getenv("DPL_TRACE_REFILL");
        const int x = e ? std::atoi(e) : kTraceRefillDefault;
        return x < 0 ? 0 : (x > kWave ? kWave : x);
    }();
    return v;
}

__global__ void KResetCounter(uint32_t* c, unsigned long long* counters) {
#pragma unroll
    for (int k = 0; k < kTraceShares; k++) c[kTraceShareStride * k] = 0u;  // one ray counter per share
    if (counters) counters[0] = counters[1] = 0ull;
}

hipError_t LaunchResetCounter(uint32_t* counter, hipStream_t s) {
    hipLaunchKernelGGL(KResetCounter, dim3(1), dim3(1), 0, s, counter, nullptr);
    return hipGetLastError();
}

int TraceGrid(int cu_count, int64_t n) {
    const int64_t need = (n + kTraceBlock - 1) / kTraceBlock;
    return int(std::max<int64_t>(1, std::min<int64_t>(int64_t(cu_count) * kTraceWgPerCu, need)));
}

// Per-lane spill entries beyond the shortest LDS ring (the SOBB kernels'); the other kernels'
// longer rings need fewer, so one size serves every trace and render kernel.
int32_t TraceSpillStride(int height) { return height >= kStackLdsSobb ? height - kStackLdsSobb + 1 : 0; }

size_t TraceSpillBytes(int cu_count, int height) {
    return size_t(cu_count) * kTraceWgPerCu * kTraceBlock * size_t(TraceSpillStride(height)) * 4;
}

int TraceGridFor(int cu_count, int64_t n, int height, size_t spill_bytes) {
    const size_t per_wg = size_t(kTraceBlock) * size_t(TraceSpillStride(height)) * 4;
    const int grid = TraceGrid(cu_count, n);
    if (per_wg == 0) return grid;
    return int(std::min<size_t>(size_t(grid), spill_bytes / per_wg));
}

// ---- coherent ray order ----
// Rays whose span inside the scene box passes the same region traverse the same nodes: tracing
// them side by side keeps a wave's lanes on common paths (fewer idle lanes per VALU
// instruction) and the nodes they read in L2. The point a quarter of the way into the span
// stands for the region (rays aimed through the scene meet surfaces early in it).
__device__ __forceinline__ uint32_t Spread10(uint32_t v) {
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// lo / hi: the scene box; q: 1023 / extent per axis (0 for a flat axis)
__global__ void KRayKeys(const float4* __restrict__ rays, int64_t n, float3 lo, float3 hi, float3 q,
                         uint32_t* __restrict__ keys, uint32_t* __restrict__ ids) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = rays[2 * i], r1 = rays[2 * i + 1];
    const float ix = 1.0f / r1.x, iy = 1.0f / r1.y, iz = 1.0f / r1.z;
    const float ax = (lo.x - r0.x) * ix, bx = (hi.x - r0.x) * ix;
    const float ay = (lo.y - r0.y) * iy, by = (hi.y - r0.y) * iy;
    const float az = (lo.z - r0.z) * iz, bz = (hi.z - r0.z) * iz;
    const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), r0.w));
    const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), r1.w));
    uint32_t key = 0xFFFFFFFFu;  // misses the box (or a NaN on both sides): last
    if (tn <= tf) {
        const float t = tn + 0.25f * (tf - tn);
        auto cell = [](float v) { return uint32_t(fminf(fmaxf(v, 0.0f), 1023.0f)); };  // NaN -> 0
        const uint32_t cx = cell((r0.x + r1.x * t - lo.x) * q.x);
        const uint32_t cy = cell((r0.y + r1.y * t - lo.y) * q.y);
        const uint32_t cz = cell((r0.z + r1.z * t - lo.z) * q.z);
        key = Spread10(cx) | (Spread10(cy) << 1) | (Spread10(cz) << 2);
    }
    keys[i] = key;
    ids[i] = uint32_t(i);
}

size_t RayOrderTempBytes(int64_t n) { return RadixSortTempBytes(n, 32, 8, false); }


hipError_t LaunchRayOrder(const dpl_ray* rays, int64_t n, const float lo[3], const float hi[3], uint32_t* buf,
                          void* temp, size_t temp_bytes, const int32_t** order, hipStream_t s) {
    *order = nullptr;
    if (n <= 0) return hipSuccess;
    if (n > int64_t(0x7FFFFFFF)) return hipErrorInvalidValue;
    uint32_t *keys = buf, *ids = buf + n, *keys_alt = buf + 2 * n, *ids_alt = buf + 3 * n;
    float q[3];
    for (int a = 0; a < 3; a++) q[a] = hi[a] > lo[a] ? 1023.0f / (hi[a] - lo[a]) : 0.0f;
    const int grid = int((n + 255) / 256);
    hipLaunchKernelGGL(KRayKeys, dim3(grid), dim3(256), 0, s, reinterpret_cast<const float4*>(rays), n,
                       make_float3(lo[0], lo[1], lo[2]), make_float3(hi[0], hi[1], hi[2]), make_float3(q[0], q[1], q[2]),
                       keys, ids);
    bool alt = false;
    const hipError_t e = LaunchRadixSort(reinterpret_cast<uint64_t*>(keys), ids, reinterpret_cast<uint64_t*>(keys_alt),
                                         ids_alt, n, 32, 8, false, temp, temp_bytes, &alt, s);
    if (e != hipSuccess) return e;
    *order = reinterpret_cast<const int32_t*>(alt ? ids_alt : ids);
    return hipGetLastError();
}

hipError_t LaunchTrace(const TraceTree& tree, const dpl_ray* rays, int64_t n, uint32_t flags, dpl_hit* hits,
                       int cu_count, uint32_t* counter, int32_t* spill, unsigned long long* counters, hipStream_t s,
                       int max_grid, const int32_t* order) {
    if (n <= 0) return hipSuccess;
    if (n > int64_t(0x7FFFFFFF)) return hipErrorInvalidValue;  // the 32-bit ray counter never wraps
    const float4* r4 = reinterpret_cast<const float4*>(rays);
    int grid = TraceGrid(cu_count, n);
    if (max_grid > 0 && grid > max_grid) grid = max_grid;
    if (grid < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(KResetCounter, dim3(1), dim3(1), 0, s, counter, counters);
    const int32_t stride = TraceSpillStride(tree.height);
#define DPL_TRACE(COUNT_, SOBB_)                                                                                   \
    hipLaunchKernelGGL((KTrace<kTraceBlock, COUNT_, SOBB_, false>), dim3(grid), dim3(kTraceBlock), 0, s, tree, r4, n, \
                       flags, hits, counter, spill, stride, counters, nullptr, nullptr, 0)
    const bool sobb = tree.ndirs > 0;
    const int refill = TraceRefillLanes();
    const int shares = TraceSharesEnv();
    const int64_t share_len = (n + shares - 1) / shares;
    if (counters) {  // the counters count the ordered traversal, as the oracle does
        if (sobb) DPL_TRACE(true, true);
        else DPL_TRACE(true, false);
    } else if (refill == 0) {
        if (sobb) DPL_TRACE(false, true);
        else DPL_TRACE(false, false);
    } else {
        const bool any = flags & DPL_TRACE_ANY, cull = flags & DPL_TRACE_CULL_BACK;
#define DPL_TRACE_R(SOBB_, ANY_)                                                                                     \
    hipLaunchKernelGGL((KTraceRefill<kTraceBlock, SOBB_, false, ANY_, kTracePending>), dim3(grid),               \
                       dim3(kTraceBlock), 0, s, tree, r4, n, cull, hits, counter, spill, stride, nullptr, nullptr, 0,  \
                       refill, order, shares, share_len)
        if (sobb) {
            if (any) DPL_TRACE_R(true, true);
            else DPL_TRACE_R(true, false);
        } else {
            if (any) DPL_TRACE_R(false, true);
            else DPL_TRACE_R(false, false);
        }
#undef DPL_TRACE_R
    }
#undef DPL_TRACE
    return hipGetLastError();
}

hipError_t LaunchTraceBatched(const FNode* nodes, const float4* tris, const BatchTree* mt, int32_t mesh_count,
                              int32_t max_height, const dpl_ray* rays, const int32_t* ray_mesh, int64_t n,
                              uint32_t flags, dpl_hit* hits, int cu_count, uint32_t* counter, int32_t* spill,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (n > int64_t(0x7FFFFFFF)) return hipErrorInvalidValue;  // the 32-bit ray counter never wraps
    hipLaunchKernelGGL(KResetCounter, dim3(1), dim3(1), 0, s, counter, nullptr);
    TraceTree tree{nodes, tris, 0, 0, max_height, 0, nullptr};
    const int grid = TraceGrid(cu_count, n);
    const int refill = TraceRefillLanes();
    const int shares = TraceSharesEnv();
    const int64_t share_len = (n + shares - 1) / shares;
    const float4* r4 = reinterpret_cast<const float4*>(rays);
    const int32_t stride = TraceSpillStride(max_height);
    if (refill == 0) {
        hipLaunchKernelGGL((KTrace<kTraceBlock, false, false, true>), dim3(grid), dim3(kTraceBlock), 0, s, tree, r4, n,
                           flags, hits, counter, spill, stride, nullptr, ray_mesh, mt, mesh_count);
    } else {
        const bool cull = flags & DPL_TRACE_CULL_BACK;
        if (flags & DPL_TRACE_ANY)
            hipLaunchKernelGGL((KTraceRefill<kTraceBlock, false, true, true, kTracePending>), dim3(grid), dim3(kTraceBlock), 0, s,
                               tree, r4, n, cull, hits, counter, spill, stride, ray_mesh, mt, mesh_count, refill,
                               nullptr, shares, share_len);
        else
            hipLaunchKernelGGL((KTraceRefill<kTraceBlock, false, true, false, kTracePending>), dim3(grid), dim3(kTraceBlock), 0, s,
                               tree, r4, n, cull, hits, counter, spill, stride, ray_mesh, mt, mesh_count, refill,
                               nullptr, shares, share_len);
    }
    return hipGetLastError();
}

}  // namespace dpl
