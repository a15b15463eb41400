// Ray traversal for gfx950: closest-hit / any-hit over the 64-byte two-child nodes, with
// the watertight ray-triangle test of Woop, Benthin and Wald (JCGT 2013) and a
// conservative slab test (Ize 2013). Each lane keeps its traversal stack in LDS (one
// column per lane, conflict-free), the rays are fetched by 64-lane waves from a global
// atomic counter (persistent grid), and a finished wave immediately takes the next batch.
//
// Closest-hit rule: smallest t with tmin < t < tmax, ties to the lower primitive id, so the
// result does not depend on the tree or on the visiting order.
#include <hip/hip_runtime.h>

#include "dpl_internal.h"

namespace dpl {

namespace {

struct WtRay {
    int kx, ky, kz;
    float Sx, Sy, Sz;
    float ox, oy, oz;
    float idx, idy, idz;
    float tmin, tmax;
};

__device__ __forceinline__ float Sel(float x, float y, float z, int k) { return k == 0 ? x : (k == 1 ? y : z); }

__device__ __forceinline__ WtRay PrepareRay(const float4 r0, const float4 r1) {
    WtRay w;
    const float dx = r1.x, dy = r1.y, dz = r1.z;
    const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    const int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    const float dkz = Sel(dx, dy, dz, kz);
    if (dkz < 0.0f) {
        const int t = kx;
        kx = ky;
        ky = t;
    }
    w.kx = kx, w.ky = ky, w.kz = kz;
    w.Sx = Sel(dx, dy, dz, kx) / dkz;
    w.Sy = Sel(dx, dy, dz, ky) / dkz;
    w.Sz = 1.0f / dkz;
    w.ox = r0.x, w.oy = r0.y, w.oz = r0.z;
    w.idx = 1.0f / dx;
    w.idy = 1.0f / dy;
    w.idz = 1.0f / dz;
    w.tmin = r0.w;
    w.tmax = r1.w;
    return w;
}

__device__ __forceinline__ bool Watertight(const WtRay& w, const float4 p0, const float4 p1, const float4 p2,
                                           bool cull, float& t, float& b1, float& b2) {
    const float A0 = p0.x - w.ox, A1 = p0.y - w.oy, A2 = p0.z - w.oz;
    const float B0 = p1.x - w.ox, B1 = p1.y - w.oy, B2 = p1.z - w.oz;
    const float C0 = p2.x - w.ox, C1 = p2.y - w.oy, C2 = p2.z - w.oz;
    const float Akz = Sel(A0, A1, A2, w.kz), Bkz = Sel(B0, B1, B2, w.kz), Ckz = Sel(C0, C1, C2, w.kz);
    const float Ax = Sel(A0, A1, A2, w.kx) - w.Sx * Akz;
    const float Ay = Sel(A0, A1, A2, w.ky) - w.Sy * Akz;
    const float Bx = Sel(B0, B1, B2, w.kx) - w.Sx * Bkz;
    const float By = Sel(B0, B1, B2, w.ky) - w.Sy * Bkz;
    const float Cx = Sel(C0, C1, C2, w.kx) - w.Sx * Ckz;
    const float Cy = Sel(C0, C1, C2, w.ky) - w.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        const double CxBy = double(Cx) * double(By), CyBx = double(Cy) * double(Bx);
        U = float(CxBy - CyBx);
        const double AxCy = double(Ax) * double(Cy), AyCx = double(Ay) * double(Cx);
        V = float(AxCy - AyCx);
        const double BxAy = double(Bx) * double(Ay), ByAx = double(By) * double(Ax);
        W = float(BxAy - ByAx);
    }
    if (cull) {
        if (U < 0.0f || V < 0.0f || W < 0.0f) return false;
    } else {
        if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    }
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = w.Sz * Akz, Bz = w.Sz * Bkz, Cz = w.Sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    if (det < 0.0f ? (T > 0.0f) : (T < 0.0f)) return false;
    t = T / det;
    b1 = V / det;
    b2 = W / det;
    return true;
}

__device__ __forceinline__ bool Box(const WtRay& w, float tfar_limit, float x0, float x1, float y0, float y1, float z0,
                                    float z1, float& tnear) {
    const float tx0 = (x0 - w.ox) * w.idx, tx1 = (x1 - w.ox) * w.idx;
    const float ty0 = (y0 - w.oy) * w.idy, ty1 = (y1 - w.oy) * w.idy;
    const float tz0 = (z0 - w.oz) * w.idz, tz1 = (z1 - w.oz) * w.idz;
    const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), w.tmin));
    const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tfar_limit));
    tnear = tn;
    return tn <= tf * 1.0000004f;
}

constexpr int kTraceBlock = 256;

}  // namespace

template <int STACK, int BLOCK>
__global__ __launch_bounds__(BLOCK) void KTrace(TraceTree tree, const float4* __restrict__ rays, int64_t n,
                                                uint32_t flags, dpl_hit* __restrict__ hits, uint32_t* counter) {
    const FNode* __restrict__ nodes = tree.nodes;
    const float4* __restrict__ tris = tree.tris;
    const int32_t node_count = tree.node_count;
    const int32_t prim_count = tree.prim_count;
    __shared__ int32_t s_stack[STACK][BLOCK];
    __shared__ uint32_t s_base[BLOCK / kWave];
    const bool any = flags & DPL_TRACE_ANY, cull = flags & DPL_TRACE_CULL_BACK;
    const int lane = LaneId(), wid = threadIdx.x / kWave;
    int32_t* const col = &s_stack[0][threadIdx.x];
    for (;;) {
        if (lane == 0) s_base[wid] = atomicAdd(counter, uint32_t(kWave));
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int64_t base = s_base[wid];
        if (base >= n) break;
        const int64_t i = base + lane;
        if (i < n) {
            const WtRay w = PrepareRay(rays[2 * i], rays[2 * i + 1]);
            float best_t = w.tmax, bb1 = 0.0f, bb2 = 0.0f;
            int32_t best = -1;
            int sp = 0;
            int32_t ref = 0;
            int64_t budget = 2 * int64_t(node_count) + 2;  // a tree visits each node at most once
            for (;;) {
                if (--budget < 0) break;
                if (ref >= node_count) break;
                if (ref >= 0) {
                    const float4* np = reinterpret_cast<const float4*>(nodes + ref);
                    const float4 a = np[0], b = np[1], c = np[2];
                    const int4 d = reinterpret_cast<const int4*>(np)[3];
                    float tl, tr;
                    const bool hl = Box(w, best_t, a.x, a.y, a.z, a.w, c.x, c.y, tl);
                    const bool hr = Box(w, best_t, b.x, b.y, b.z, b.w, c.z, c.w, tr);
                    if (hl && hr) {
                        const bool rfirst = tr < tl;
                        col[sp * BLOCK] = rfirst ? d.x : d.y;
                        sp++;
                        ref = rfirst ? d.y : d.x;
                    } else if (hl) {
                        ref = d.x;
                    } else if (hr) {
                        ref = d.y;
                    } else {
                        if (sp == 0) break;
                        ref = col[--sp * BLOCK];
                    }
                } else {
                    const int32_t packed = ~ref;
                    const int32_t b0 = packed >> 4;
                    const int32_t cnt = b0 + (packed & 15) <= prim_count ? (packed & 15) : 0;
                    bool done = false;
                    for (int32_t k = b0; k < b0 + cnt; k++) {
                        const float4 p0 = tris[3 * int64_t(k)], p1 = tris[3 * int64_t(k) + 1],
                                     p2 = tris[3 * int64_t(k) + 2];
                        float t, u1, u2;
                        if (!Watertight(w, p0, p1, p2, cull, t, u1, u2)) continue;
                        if (!(t > w.tmin)) continue;
                        const int32_t prim = __float_as_int(p0.w);
                        if (any) {
                            if (t < w.tmax) {
                                best_t = t, bb1 = u1, bb2 = u2, best = prim;
                                done = true;
                                break;
                            }
                            continue;
                        }
                        if (t < best_t || (t == best_t && best >= 0 && prim < best)) {
                            best_t = t, bb1 = u1, bb2 = u2, best = prim;
                        }
                    }
                    if (done || sp == 0) break;
                    ref = col[--sp * BLOCK];
                }
            }
            dpl_hit h;
            h.t = best >= 0 ? best_t : 0.0f;
            h.b1 = best >= 0 ? bb1 : 0.0f;
            h.b2 = best >= 0 ? bb2 : 0.0f;
            h.prim = best;
            hits[i] = h;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

__global__ void KResetCounter(uint32_t* c) { c[0] = 0u; }

hipError_t LaunchTrace(const TraceTree& tree, const dpl_ray* rays, int64_t n, uint32_t flags, dpl_hit* hits,
                       int cu_count, uint32_t* counter, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (n > int64_t(0xFFFFFFFFu) - 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(KResetCounter, dim3(1), dim3(1), 0, s, counter);
    const float4* r4 = reinterpret_cast<const float4*>(rays);
    const int64_t waves_needed = (n + kWave - 1) / kWave;
    auto grid_for = [&](int block, int per_cu) {
        int64_t g = int64_t(cu_count) * per_cu;
        int64_t need = (waves_needed * kWave + block - 1) / block;
        return int(std::max<int64_t>(1, std::min(g, need)));
    };
    const int h = tree.height;
    if (h <= 32) {
        hipLaunchKernelGGL((KTrace<32, 256>), dim3(grid_for(256, 8)), dim3(256), 0, s, tree, r4, n, flags, hits,
                           counter);
    } else if (h <= 64) {
        hipLaunchKernelGGL((KTrace<64, 256>), dim3(grid_for(256, 4)), dim3(256), 0, s, tree, r4, n, flags, hits,
                           counter);
    } else if (h <= 128) {
        hipLaunchKernelGGL((KTrace<128, 128>), dim3(grid_for(128, 8)), dim3(128), 0, s, tree, r4, n, flags, hits,
                           counter);
    } else if (h <= 256) {
        hipLaunchKernelGGL((KTrace<256, 64>), dim3(grid_for(64, 16)), dim3(64), 0, s, tree, r4, n, flags, hits,
                           counter);
    } else {
        return hipErrorNotSupported;
    }
    return hipGetLastError();
}

}  // namespace dpl
