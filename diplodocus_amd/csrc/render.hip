// Differentiable diffuse path tracing for gfx950.
//
// Forward: one lane per pixel (64-pixel batches per wave from a global counter), spp paths
// per pixel in order, each path up to max_depth closest-hit traversals over the BVH;
// emission is collected on every hit, bounces are cosine-sampled around the geometric
// normal facing the incoming ray, throughput is multiplied by the hit material's albedo.
// Random numbers: PCG32 (O'Neill 2014), one stream per (pixel, sample).
//
// Adjoint by path replay (after Vicini, Speierer, Jakob 2021): the path is re-traced with
// the same random numbers (the albedo drives no sampling decision, so the replayed path is
// the forward path), its vertex materials and throughputs kept in registers, and the
// backward recurrence  S_{k-1} = E_{m_k} + rho_{m_k} S_k  gives
//   dL/dE_{m_k}   += delta * beta_k
//   dL/drho_{m_k} += delta * beta_k * S_k,        delta = dL/dI(pixel) / spp,
// accumulated with LDS float atomics per workgroup and flushed once per workgroup.
#include <hip/hip_runtime.h>

#include "dpl_internal.h"
#include "traverse.h"
#include "shade_common.h"

namespace dpl {

namespace {

constexpr int kMaxMaterials = 256;
constexpr float kRayEps = 1e-6f;

struct Pcg {
    uint64_t state, inc;
};

__device__ __forceinline__ uint32_t PcgNext(Pcg& r) {
    const uint64_t old = r.state;
    r.state = old * 6364136223846793005ull + r.inc;
    const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

__device__ __forceinline__ void PcgSeed(Pcg& r, uint64_t seed, uint64_t stream) {
    r.state = 0u;
    r.inc = (stream << 1u) | 1u;
    PcgNext(r);
    r.state += seed;
    PcgNext(r);
}

__device__ __forceinline__ float PcgFloat(Pcg& r) { return float(PcgNext(r) >> 8) * (1.0f / 16777216.0f); }

// sin / cos of 2*pi*u: quadrant reduction and Taylor polynomials (degree 11 / 12) on [0, pi/2)
__device__ __forceinline__ void SinCos2Pi(float u, float& s, float& c) {
    const float x = u * 4.0f;
    int q = int(x);
    if (q > 3) q = 3;
    const float f = x - float(q);
    const float th = f * 1.57079637f;
    const float t2 = th * th;
    float sp = -2.50521084e-8f;
    sp = sp * t2 + 2.75573192e-6f;
    sp = sp * t2 + -1.98412698e-4f;
    sp = sp * t2 + 8.33333333e-3f;
    sp = sp * t2 + -1.66666667e-1f;
    sp = sp * t2 + 1.0f;
    sp = sp * th;
    float cp = 2.08767570e-9f;
    cp = cp * t2 + -2.75573192e-7f;
    cp = cp * t2 + 2.48015873e-5f;
    cp = cp * t2 + -1.38888889e-3f;
    cp = cp * t2 + 4.16666667e-2f;
    cp = cp * t2 + -0.5f;
    cp = cp * t2 + 1.0f;
    s = q == 0 ? sp : (q == 1 ? cp : (q == 2 ? -sp : -cp));
    c = q == 0 ? cp : (q == 1 ? -sp : (q == 2 ? -cp : sp));
}

struct RenderArgs {
    TraceTree tree;
    const int32_t* mat_id;
    const float* albedo;
    const float* emission;
    int32_t M;
    int32_t W, H, spp, depth;
    uint64_t seed;
    float basis[12];  // pos, p00, qw, qh
    float far;
    const float* d_image;
    float* image;
    float* grad_albedo;
    float* grad_emission;
    int32_t* spill;
    int32_t spill_stride;
};

constexpr int kRenderStack = 16;

}  // namespace

// A sample per lane: spp a power of two that divides the wave, and a 31-bit item count.
__host__ __device__ inline bool SampleLanes(int spp, int64_t pixels) {
    return spp > 1 && spp <= kWave && (spp & (spp - 1)) == 0 && pixels * spp < (int64_t(1) << 31);
}

template <bool ADJ, bool SOBB, int BLOCK, int MAXD>
__global__ __launch_bounds__(BLOCK) void KRender(RenderArgs a, uint32_t* counter) {
    __shared__ int32_t s_stack[kRenderStack][BLOCK];
    __shared__ uint32_t s_base[BLOCK / kWave];
    __shared__ float s_grad[ADJ ? 6 * kMaxMaterials : 1];
    __shared__ float s_idn[SOBB ? kMaxDirs : 1][BLOCK];
    __shared__ float s_dirs[SOBB ? 3 * kMaxDirs : 1];
    if (SOBB) {
        for (int k = threadIdx.x; k < 3 * a.tree.ndirs; k += BLOCK) s_dirs[k] = a.tree.dirs[k];
        __syncthreads();
    }
    const int lane = LaneId(), wid = threadIdx.x / kWave;
    LaneStack<kRenderStack, BLOCK> st{&s_stack[0][threadIdx.x],
                                      a.spill + (int64_t(blockIdx.x) * BLOCK + threadIdx.x) * a.spill_stride, 0};
    SobbRay sr{s_dirs, &s_idn[0][threadIdx.x], 0.0f, 0.0f, 0.0f, 0.0f};
    TraceCounters cnt{0u, 0u};
    if (ADJ) {
        for (int i = threadIdx.x; i < 6 * a.M; i += BLOCK) s_grad[i] = 0.0f;
        __syncthreads();
    }
    const int64_t pixels = int64_t(a.W) * a.H;
    // One sample per lane when spp divides the wave (a pixel's samples then sit in adjacent
    // lanes of one wave, spp lanes per pixel): spp times the parallelism of a pixel per lane.
    // The forward sums a pixel's samples in sample order through shuffles, the same additions
    // as one lane's loop; the adjoint's per-sample terms are independent.
    const bool spl = SampleLanes(a.spp, pixels);
    const int sh = spl ? __builtin_ctz(unsigned(a.spp)) : 0;
    const int64_t items = pixels << sh;
    const V3 pos{a.basis[0], a.basis[1], a.basis[2]}, p00{a.basis[3], a.basis[4], a.basis[5]};
    const V3 qw{a.basis[6], a.basis[7], a.basis[8]}, qh{a.basis[9], a.basis[10], a.basis[11]};
    for (;;) {
        if (lane == 0) s_base[wid] = atomicAdd(counter, uint32_t(kWave));
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int64_t base = s_base[wid];
        if (base >= items) break;
        const int64_t q = base + lane;
        const int64_t p = q >> sh;
        const int s0 = spl ? int(q & (a.spp - 1)) : 0, s1 = spl ? s0 + 1 : a.spp;
        V3 acc{0.0f, 0.0f, 0.0f};
        if (q < items) {
            const int x = int(p % a.W), y = int(p / a.W);
            V3 delta{0.0f, 0.0f, 0.0f};
            if (ADJ) {
                const float inv = float(a.spp);
                delta = {a.d_image[3 * p] / inv, a.d_image[3 * p + 1] / inv, a.d_image[3 * p + 2] / inv};
            }
            for (int s = s0; s < s1; s++) {
                Pcg rng;
                PcgSeed(rng, a.seed, uint64_t(p) * uint64_t(a.spp) + uint64_t(s));
                float fx = float(x), fy = float(y);
                if (a.spp > 1) {
                    const float j1 = PcgFloat(rng);
                    const float j2 = PcgFloat(rng);
                    fx = fx + (j1 - 0.5f);
                    fy = fy + (j2 - 0.5f);
                }
                V3 d = Normalize(Add(Add(p00, Mul(qw, fx)), Mul(qh, fy)));
                V3 o = pos;
                V3 beta{1.0f, 1.0f, 1.0f};
                V3 L{0.0f, 0.0f, 0.0f};
                int32_t mats[MAXD];
                V3 betas[MAXD];
                int K = 0;
                for (int dep = 0; dep < a.depth; dep++) {
                    const WtRay w = PrepareRay(o.x, o.y, o.z, kRayEps, d.x, d.y, d.z, a.far);
                    if (SOBB) PrepareSobb<BLOCK>(sr, s_dirs, a.tree.ndirs, o.x, o.y, o.z, d.x, d.y, d.z);
                    const HitResult h = Traverse<false, SOBB>(a.tree, w, sr, false, st, cnt);
                    if (h.prim < 0) break;
                    const int32_t m = a.mat_id[h.prim];
                    if (ADJ) {
#pragma unroll
                        for (int j = 0; j < MAXD; j++) {
                            if (j == K) {
                                mats[j] = m;
                                betas[j] = beta;
                            }
                        }
                    } else {
                        const V3 E{a.emission[3 * m], a.emission[3 * m + 1], a.emission[3 * m + 2]};
                        L = Add(L, Mul(beta, E));
                    }
                    K++;
                    if (dep + 1 == a.depth) break;
                    const float4* tv = a.tree.tris + 3 * int64_t(h.slot);
                    const float4 q0 = tv[0], q1 = tv[1], q2 = tv[2];
                    const V3 v0{q0.x, q0.y, q0.z}, v1{q1.x, q1.y, q1.z}, v2{q2.x, q2.y, q2.z};
                    V3 nrm = Normalize(Cross(Sub(v1, v0), Sub(v2, v0)));
                    if (Dot(nrm, d) > 0.0f) nrm = {-nrm.x, -nrm.y, -nrm.z};
                    const V3 ph = Add(o, Mul(d, h.t));
                    const float eps = RayEpsilon(ph, h.t);
                    o = Add(ph, Mul(nrm, eps));
                    const float u1 = PcgFloat(rng);
                    const float u2 = PcgFloat(rng);
                    const float rr = sqrtf(u1);
                    float sn, cs;
                    SinCos2Pi(u2, sn, cs);
                    const float lx = rr * cs, ly = rr * sn, lz = sqrtf(Max2(0.0f, 1.0f - u1));
                    const float sign = nrm.z >= 0.0f ? 1.0f : -1.0f;
                    const float aa = -1.0f / (sign + nrm.z);
                    const float bb = nrm.x * nrm.y * aa;
                    const V3 t1{1.0f + sign * nrm.x * nrm.x * aa, sign * bb, -sign * nrm.x};
                    const V3 t2{bb, sign + nrm.y * nrm.y * aa, -nrm.y};
                    d = Add(Add(Mul(t1, lx), Mul(t2, ly)), Mul(nrm, lz));
                    beta = Mul(beta, V3{a.albedo[3 * m], a.albedo[3 * m + 1], a.albedo[3 * m + 2]});
                }
                if (ADJ) {
                    V3 S{0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int k = MAXD - 1; k >= 0; k--) {
                        if (k < K) {
                            const int32_t m = mats[k];
                            const V3 b = betas[k];
                            float* g = &s_grad[6 * m];
                            atomicAdd(&g[3], delta.x * b.x);
                            atomicAdd(&g[4], delta.y * b.y);
                            atomicAdd(&g[5], delta.z * b.z);
                            atomicAdd(&g[0], delta.x * b.x * S.x);
                            atomicAdd(&g[1], delta.y * b.y * S.y);
                            atomicAdd(&g[2], delta.z * b.z * S.z);
                            const V3 E{a.emission[3 * m], a.emission[3 * m + 1], a.emission[3 * m + 2]};
                            const V3 R{a.albedo[3 * m], a.albedo[3 * m + 1], a.albedo[3 * m + 2]};
                            S = Add(E, Mul(R, S));
                        }
                    }
                } else {
                    acc = Add(acc, L);
                }
            }
        }
        if (!ADJ) {
            if (spl) {  // the pixel's first lane adds the others' samples in sample order
                V3 tot = acc;
                for (int k = 1; k < a.spp; k++) {
                    const int src = lane + k < kWave ? lane + k : lane;
                    const V3 o{__shfl(acc.x, src, kWave), __shfl(acc.y, src, kWave), __shfl(acc.z, src, kWave)};
                    tot = Add(tot, o);
                }
                acc = tot;
            }
            if (q < items && (!spl || (q & (a.spp - 1)) == 0)) {
                const float inv = float(a.spp);
                a.image[3 * p] = acc.x / inv;
                a.image[3 * p + 1] = acc.y / inv;
                a.image[3 * p + 2] = acc.z / inv;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (ADJ) {
        __syncthreads();
        for (int i = threadIdx.x; i < 3 * a.M; i += BLOCK) {
            const int m = i / 3, c = i % 3;
            const float ga = s_grad[6 * m + c], ge = s_grad[6 * m + 3 + c];
            if (ga != 0.0f) atomicAdd(&a.grad_albedo[i], ga);
            if (ge != 0.0f) atomicAdd(&a.grad_emission[i], ge);
        }
    }
}

__global__ void KZero(float* p, int64_t n) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        p[i] = 0.0f;
}

hipError_t LaunchResetCounter(uint32_t* counter, hipStream_t s);

int TraceGrid(int cu_count, int64_t n);
int32_t TraceSpillStride(int height);

hipError_t LaunchRender(bool adjoint, const TraceTree& tree, const dpl_render_params& rp, const float* basis,
                        const int32_t* mat_id, const float* albedo, const float* emission, const float* d_image,
                        float* image, float* grad_albedo, float* grad_emission, int cu_count, uint32_t* counter,
                        int32_t* spill, hipStream_t s) {
    RenderArgs a{};
    a.tree = tree;
    a.mat_id = mat_id;
    a.albedo = albedo;
    a.emission = emission;
    a.M = rp.material_count;
    a.W = rp.width, a.H = rp.height, a.spp = rp.spp, a.depth = rp.max_depth;
    a.seed = rp.seed;
    for (int k = 0; k < 12; k++) a.basis[k] = basis[k];
    a.far = rp.camera[10];
    a.d_image = d_image;
    a.image = image;
    a.grad_albedo = grad_albedo;
    a.grad_emission = grad_emission;
    a.spill = spill;
    a.spill_stride = TraceSpillStride(tree.height);
    if (a.M < 1 || a.M > kMaxMaterials) return hipErrorInvalidValue;
    if (a.depth < 1 || (adjoint && a.depth > 16)) return hipErrorInvalidValue;
    hipError_t e = LaunchResetCounter(counter, s);
    if (e != hipSuccess) return e;
    if (adjoint) {
        hipLaunchKernelGGL(KZero, dim3(1), dim3(256), 0, s, grad_albedo, int64_t(3 * a.M));
        hipLaunchKernelGGL(KZero, dim3(1), dim3(256), 0, s, grad_emission, int64_t(3 * a.M));
    }
    constexpr int B = 256;  // same block and grid policy as the trace kernel (shared spill sizing)
    const int64_t pixels = int64_t(a.W) * a.H;
    const int grid = TraceGrid(cu_count, SampleLanes(a.spp, pixels) ? pixels * a.spp : pixels);
#define DPL_RENDER(ADJ_, SOBB_, MD_) \
    hipLaunchKernelGGL((KRender<ADJ_, SOBB_, B, MD_>), dim3(grid), dim3(B), 0, s, a, counter)
    const bool sobb = tree.ndirs > 0;
    if (!adjoint) {
        if (sobb) DPL_RENDER(false, true, 1);
        else DPL_RENDER(false, false, 1);
    } else if (a.depth <= 4) {
        if (sobb) DPL_RENDER(true, true, 4);
        else DPL_RENDER(true, false, 4);
    } else if (a.depth <= 8) {
        if (sobb) DPL_RENDER(true, true, 8);
        else DPL_RENDER(true, false, 8);
    } else {
        if (sobb) DPL_RENDER(true, true, 16);
        else DPL_RENDER(true, false, 16);
    }
#undef DPL_RENDER
    return hipGetLastError();
}

}  // namespace dpl
