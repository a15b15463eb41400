// Whitted ray tracing over the BVH for gfx950 — the reference's CpuRaytracer
// (src/renderer/cpu_raytracer.cc:50-288) on the GPU: per pixel a primary ray from the
// reference's camera basis; at each hit either the emission of an emissive material, or
// area-light next-event estimation (hash-sampled light points, seed 42, shadow rays with the
// geometric-normal offset) with Phong shading, then mirror reflection and refraction (total
// internal reflection falls back to reflection) to max_depth.
//
// The recursion I = L_local + I(reflected) ks + I(refracted) kt is evaluated as a weighted
// sum over the ray tree, depth first, reflection subtree first: a per-lane stack of
// (ray, weight, depth) items. Persistent grid: waves take 64 / G pixels at a time from a
// counter, G lanes per pixel (1-8, more for small images), which trace the pixel's rays in
// step and split each hit's light samples.
// Every float operation follows the oracle's restatement in order (no FMA contraction), so
// the image is bit-identical to it. With COUNT the kernel also keeps the reference's ray
// counters (primary, secondary, shadow: cpu_raytracer.cc:103,131,147,223) and the traversal
// work of every query, summed per wave and added to six 64-bit totals.
#include <hip/hip_runtime.h>

#include "dpl_internal.h"
#include "shade_common.h"
#include "traverse.h"

namespace dpl {

namespace {

constexpr float kEps = 1e-6f;  // the reference's kEpsilon (src/util/util.h:15)
constexpr int kWhittedStack = 16;
constexpr int kWhittedLanesMax = 8;  // most lanes per pixel (light samples traced side by side)

// x^y for x >= 0, y >= 0 by exp2(y log2 x): exponent extraction, the atanh series of the
// mantissa for log2, the Taylor series of e^(f ln 2) for the fraction (same as the oracle).
__device__ __forceinline__ float PowPos(float x, float y) {
    if (y == 0.0f) return 1.0f;
    if (!(x >= 1.17549435e-38f)) return 0.0f;
    const uint32_t bits = __float_as_uint(x);
    int e = int(bits >> 23) - 127;
    float m = __uint_as_float((bits & 0x7FFFFFu) | 0x3F800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e = e + 1;
    }
    const float t = (m - 1.0f) / (m + 1.0f);
    const float t2 = t * t;
    float p = 0.111111111f;
    p = p * t2 + 0.142857143f;
    p = p * t2 + 0.2f;
    p = p * t2 + 0.333333333f;
    p = p * t2 + 1.0f;
    const float lg = float(e) + (p * t) * 2.88539008f;
    const float z = y * lg;
    if (z < -126.0f) return 0.0f;
    if (z >= 128.0f) return __uint_as_float(0x7F800000u);
    const float fl = floorf(z);
    const int n = int(fl);
    const float g = (z - fl) * 0.693147181f;
    float q = 2.75573192e-6f;
    q = q * g + 2.48015873e-5f;
    q = q * g + 1.98412698e-4f;
    q = q * g + 1.38888889e-3f;
    q = q * g + 8.33333333e-3f;
    q = q * g + 4.16666667e-2f;
    q = q * g + 1.66666667e-1f;
    q = q * g + 0.5f;
    q = q * g + 1.0f;
    q = q * g + 1.0f;
    return q * __uint_as_float(uint32_t(n + 127) << 23);
}

__device__ __forceinline__ uint32_t HashU32(uint32_t x) {  // src/util/util.h:24-32
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// cpu_raytracer.cc:28-39 with U01FromU32 (util.h:34-37)
__device__ __forceinline__ float AreaLightSample01(uint32_t px, uint32_t py, uint32_t light, uint32_t s, uint32_t dim) {
    uint32_t key = 0;
    key ^= 42u * 0x9e3779b1U;
    key ^= px * 0x165667b1U;
    key ^= py * 0xd3a2646cU;
    key ^= light * 0x85ebca6bU;
    key ^= s * 0xc2b2ae35U;
    key ^= dim * 0x27d4eb2fU;
    return float(HashU32(key) >> 8) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ V3 Neg(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 DivS(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ bool NonZero(const float* a) { return a[0] != 0.0f || a[1] != 0.0f || a[2] != 0.0f; }
__device__ __forceinline__ V3 Ld(const float* a) { return {a[0], a[1], a[2]}; }
__device__ __forceinline__ float Len(V3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }

struct WhittedArgs {
    TraceTree tree;
    const float* tri;  // n x 9, original order
    const float* vn;   // n x 9 vertex normals (all-zero rows: none) or null
    const int32_t* mat_id;
    const dpl_material* mats;
    const int32_t* lights;  // emissive triangles, in triangle order
    int32_t light_count;
    int32_t W, H, max_depth, samples, cull;
    float background[3];
    float basis[12];  // pos, p00, qw, qh
    float far;
    float* image;
    int32_t* spill;
    int32_t spill_stride;
};

struct WHit {
    V3 pos, normal, gn;
    float t, eps;
    int32_t prim;
};

struct Item {
    V3 o, d, w;
    float tmax;
    int depth;
};

template <bool SOBB, int BLOCK>
struct Tracer {
    const WhittedArgs& a;
    LaneStack<kWhittedStack, BLOCK>& st;
    SobbRay& sr;
    const float* s_dirs;
    TraceCounters cnt{0u, 0u};
    uint32_t queries = 0;  // closest-hit and any-hit traces
    uint32_t shadow = 0;   // IsShadowed calls

    // count: this lane's trace is one of the reference's queries (a lane group traces its
    // pixel's closest-hit rays on every lane; only the group's first lane counts them)
    __device__ HitResult Trace(V3 o, V3 d, float tmax, bool any, bool cull, bool count = true) {
        const WtRay w = PrepareRay(o.x, o.y, o.z, kEps, d.x, d.y, d.z, tmax);
        if (SOBB) PrepareSobb<BLOCK>(sr, s_dirs, a.tree.ndirs, o.x, o.y, o.z, d.x, d.y, d.z);
        TraceCounters none{0u, 0u};
        TraceCounters& c = count ? cnt : none;
        queries += count ? 1u : 0u;
        return any ? Traverse<true, SOBB>(a.tree, w, sr, cull, st, c) : Traverse<false, SOBB>(a.tree, w, sr, cull, st, c);
    }
};

__device__ __forceinline__ V3 GeomNormal(const float* t) {
    const V3 v0 = Ld(t), v1 = Ld(t + 3), v2 = Ld(t + 6);
    return Normalize(Cross(Sub(v1, v0), Sub(v2, v0)));
}

__device__ __forceinline__ V3 OffsetRayOrigin(V3 pos, float eps, V3 gn, V3 dir) {
    const V3 off = Dot(dir, gn) > 0.0f ? gn : Neg(gn);
    return Add(pos, Mul(off, eps));
}

// Area-light next-event estimation of one hit (cpu_raytracer.cc:135-182) by a group of G
// adjacent lanes that hold the same hit: lane g takes samples g, g + G, ... of each light, and
// after every round of G samples the group adds the round's contributions to the colour in
// sample order (shuffles from the group's lanes), so every lane ends with the colour of the
// reference's sequential loop, bit for bit. Shadow rays are counted by the lane that traces
// them. All lanes of a group run this together (their control flow is the same up to here).
template <bool SOBB, int BLOCK, int G>
__device__ V3 LocalAreaLights(const WhittedArgs& a, Tracer<SOBB, BLOCK>& T, int x, int y, const WHit& h,
                              const dpl_material& m, V3 cam, int g, int lead) {
    V3 color{0.0f, 0.0f, 0.0f};
    const int cnt = a.samples;
    for (int al = 0; al < a.light_count; al++) {
        const int32_t lt = a.lights[al];
        const float* ltri = a.tri + 9 * int64_t(lt);
        const V3 lcolor = Ld(a.mats[a.mat_id[lt]].emission);
        if (lt == h.prim) return lcolor;
        const V3 v0 = Ld(ltri), v1 = Ld(ltri + 3), v2 = Ld(ltri + 6);
        const float area = 0.5f * Len(Cross(Sub(v1, v0), Sub(v2, v0)));
        const V3 lgn = Normalize(Cross(Sub(v1, v0), Sub(v2, v0)));
        for (int r0 = 0; r0 < cnt; r0 += G) {
            const int s = r0 + g;
            V3 c{0.0f, 0.0f, 0.0f};
            int lit = 0;
            if (s < cnt) {
                const float r1 = AreaLightSample01(uint32_t(x), uint32_t(y), uint32_t(al), uint32_t(s), 0u);
                const float r2 = AreaLightSample01(uint32_t(x), uint32_t(y), uint32_t(al), uint32_t(s), 1u);
                const float u = r1 + r2 > 1.0f ? 1.0f - r1 : r1;
                const float v = r1 + r2 > 1.0f ? 1.0f - r2 : r2;
                const V3 lpos = Add(Add(v0, Mul(Sub(v1, v0), u)), Mul(Sub(v2, v0), v));
                // IsShadowed (cpu_raytracer.cc:222-236), counted on entry
                T.shadow++;
                const V3 to_l = Sub(lpos, h.pos);
                const float dist = Len(to_l);
                bool shadowed = false;
                if (!(dist < kEps)) {
                    const V3 sd = DivS(to_l, dist);
                    const V3 so = OffsetRayOrigin(h.pos, h.eps, h.gn, sd);
                    shadowed = T.Trace(so, sd, dist * (1.0f - h.eps), true, false).prim >= 0;
                }
                if (!shadowed) {
                    const float d = Len(to_l);
                    const V3 d_l = DivS(to_l, d);
                    const float w = ((area * Max2(0.0f, Dot(lgn, Neg(d_l)))) * Max2(0.0f, Dot(h.normal, d_l))) /
                                    ((float(cnt) * d * d) + kEps);
                    const V3 lc = Mul(lcolor, w);
                    // Phong (cpu_raytracer.cc:231-255)
                    const V3 pl = Normalize(Sub(lpos, h.pos));
                    const V3 pv = Normalize(Sub(cam, h.pos));
                    const V3 pr = Normalize(Sub(Mul(Mul(h.normal, 2.0f), Dot(h.normal, pl)), pl));
                    const V3 I_a = Mul(lc, V3{0.0f, 0.0f, 0.0f});
                    const V3 I_d = Mul(Mul(lc, Ld(m.diffuse)), Max2(0.0f, Dot(h.normal, pl)));
                    const V3 I_s = Mul(Mul(lc, Ld(m.specular)), PowPos(Max2(0.0f, Dot(pv, pr)), m.shininess));
                    c = Add(Add(I_a, I_d), I_s);
                    lit = 1;
                }
            }
            // the round's samples in order (a shadowed or missing sample adds nothing)
            if constexpr (G == 1) {
                if (lit) color = Add(color, c);
            } else {
#pragma unroll
                for (int k = 0; k < G; k++) {
                    const int src = lead + k;
                    const V3 o{__shfl(c.x, src, kWave), __shfl(c.y, src, kWave), __shfl(c.z, src, kWave)};
                    if (__shfl(lit, src, kWave)) color = Add(color, o);
                }
            }
        }
    }
    return color;
}

}  // namespace

template <bool SOBB, int BLOCK, bool COUNT, int G>
__global__ __launch_bounds__(BLOCK) void KWhitted(WhittedArgs a, uint32_t* counter, unsigned long long* totals) {
    __shared__ int32_t s_stack[kWhittedStack][BLOCK];
    __shared__ uint32_t s_base[BLOCK / kWave];
    __shared__ float s_idn[SOBB ? kMaxDirs : 1][BLOCK];
    __shared__ float s_dirs[SOBB ? 3 * kMaxDirs : 1];
    if (SOBB) {
        for (int k = threadIdx.x; k < 3 * a.tree.ndirs; k += BLOCK) s_dirs[k] = a.tree.dirs[k];
        __syncthreads();
    }
    const int lane = LaneId(), wid = threadIdx.x / kWave;
    LaneStack<kWhittedStack, BLOCK> st{&s_stack[0][threadIdx.x],
                                       a.spill + (int64_t(blockIdx.x) * BLOCK + threadIdx.x) * a.spill_stride, 0};
    SobbRay sr{s_dirs, &s_idn[0][threadIdx.x], 0.0f, 0.0f, 0.0f, 0.0f};
    Tracer<SOBB, BLOCK> T{a, st, sr, s_dirs};
    const int64_t pixels = int64_t(a.W) * a.H;
    const V3 cam{a.basis[0], a.basis[1], a.basis[2]}, p00{a.basis[3], a.basis[4], a.basis[5]};
    const V3 qw{a.basis[6], a.basis[7], a.basis[8]}, qh{a.basis[9], a.basis[10], a.basis[11]};
    const V3 bg{a.background[0], a.background[1], a.background[2]};
    uint32_t primary = 0, secondary = 0;
    // G adjacent lanes per pixel: each traces the pixel's rays (the same ones, in step) and a
    // share of every hit's light samples
    const int g = lane % G, lead = lane - g;
    const bool first = g == 0;
    for (;;) {
        if (lane == 0) s_base[wid] = atomicAdd(counter, uint32_t(kWave / G));
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int64_t base = s_base[wid];
        if (base >= pixels) break;
        const int64_t p = base + lane / G;
        if (p < pixels) {
            primary += first ? 1u : 0u;
            const int x = int(p % a.W), y = int(p / a.W);
            Item stack[kDplWhittedMaxDepth + 2];
            int sp = 0;
            stack[sp++] = Item{cam, Normalize(Add(Add(p00, Mul(qw, float(x))), Mul(qh, float(y)))),
                               V3{1.0f, 1.0f, 1.0f}, a.far, 0};
            V3 acc{0.0f, 0.0f, 0.0f};
            while (sp > 0) {
                const Item it = stack[--sp];
                const HitResult hr = T.Trace(it.o, it.d, it.tmax, false, a.cull != 0, first);
                if (hr.prim < 0) {
                    acc = Add(acc, Mul(it.w, bg));
                    continue;
                }
                const dpl_material& m = a.mats[a.mat_id[hr.prim]];
                if (NonZero(m.emission)) {
                    acc = Add(acc, Mul(it.w, Ld(m.emission)));
                    continue;
                }
                WHit h;
                h.t = hr.t;
                h.prim = hr.prim;
                h.pos = Add(it.o, Mul(it.d, hr.t));
                const float b0 = (1.0f - hr.b1) - hr.b2;
                h.eps = RayEpsilon(h.pos, h.t);
                h.gn = GeomNormal(a.tri + 9 * int64_t(hr.prim));
                h.normal = h.gn;
                if (a.vn) {
                    const float* vn = a.vn + 9 * int64_t(hr.prim);
                    if (NonZero(vn) || NonZero(vn + 3) || NonZero(vn + 6))
                        h.normal = Normalize(Add(Add(Mul(Ld(vn), b0), Mul(Ld(vn + 3), hr.b1)), Mul(Ld(vn + 6), hr.b2)));
                }
                acc = Add(acc, Mul(it.w, LocalAreaLights<SOBB, BLOCK, G>(a, T, x, y, h, m, cam, g, lead)));
                if (it.depth < a.max_depth) {
                    if (NonZero(m.transmittance)) {  // cpu_raytracer.cc:264-287
                        const V3 v = Neg(it.d);
                        V3 n = h.normal;
                        float eta = m.ior;
                        float cos_i = Dot(v, h.normal);
                        if (cos_i < 0.0f) {
                            eta = m.r_ior;
                            cos_i = -cos_i;
                            n = Neg(n);
                        }
                        const float sin2_i = Max2(0.0f, 1.0f - (cos_i * cos_i));
                        const float sin2_t = sin2_i / (eta * eta);
                        V3 dir;
                        if (sin2_t >= 1.0f) {
                            dir = Add(Neg(v), Mul(h.normal, 2.0f * Dot(v, h.normal)));
                        } else {
                            const float cos_t = sqrtf(1.0f - sin2_t);
                            dir = Add(DivS(Neg(v), eta), Mul(n, cos_i / eta - cos_t));
                        }
                        stack[sp++] = Item{OffsetRayOrigin(h.pos, h.eps, h.gn, dir), dir, Mul(it.w, Ld(m.transmittance)),
                                           it.tmax, it.depth + 1};
                        secondary += first ? 1u : 0u;
                    }
                    if (NonZero(m.specular)) {  // cpu_raytracer.cc:257-262
                        const V3 dv = Neg(it.d);
                        const V3 dir = Add(Neg(dv), Mul(h.normal, 2.0f * Dot(dv, h.normal)));
                        stack[sp++] = Item{OffsetRayOrigin(h.pos, h.eps, h.gn, dir), dir, Mul(it.w, Ld(m.specular)),
                                           it.tmax, it.depth + 1};
                        secondary += first ? 1u : 0u;
                    }
                }
            }
            if (first) {
                a.image[3 * p] = acc.x;
                a.image[3 * p + 1] = acc.y;
                a.image[3 * p + 2] = acc.z;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (COUNT) {
        unsigned long long v[6] = {primary, secondary, T.shadow, T.queries, T.cnt.nodes, T.cnt.tris};
#pragma unroll
        for (int k = 0; k < 6; k++) {
#pragma unroll
            for (int d = kWave / 2; d > 0; d >>= 1) v[k] += __shfl_xor(v[k], d, kWave);
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 6; k++) atomicAdd(&totals[k], v[k]);
        }
    }
}

__global__ void KResetWhitted(uint32_t* counter, unsigned long long* totals) {
    counter[0] = 0u;
    if (totals)
        for (int k = 0; k < 6; k++) totals[k] = 0ull;
}

int TraceGrid(int cu_count, int64_t n);
int32_t TraceSpillStride(int height);

hipError_t LaunchWhitted(const TraceTree& tree, const dpl_whitted_params& wp, const float* basis, const float* tri,
                         const float* vn, const int32_t* mat_id, const dpl_material* mats, const int32_t* lights,
                         int32_t light_count, float* image, int cu_count, uint32_t* counter, int32_t* spill,
                         unsigned long long* totals, hipStream_t s) {
    WhittedArgs a{};
    a.tree = tree;
    a.tri = tri;
    a.vn = vn;
    a.mat_id = mat_id;
    a.mats = mats;
    a.lights = lights;
    a.light_count = light_count;
    a.W = wp.width, a.H = wp.height, a.max_depth = wp.max_depth, a.samples = wp.area_light_sample_cnt;
    a.cull = wp.backface_culling;
    for (int k = 0; k < 3; k++) a.background[k] = wp.background[k];
    for (int k = 0; k < 12; k++) a.basis[k] = basis[k];
    a.far = wp.camera[10];
    a.image = image;
    a.spill = spill;
    a.spill_stride = TraceSpillStride(tree.height);
    if (a.max_depth < 0 || a.max_depth > kDplWhittedMaxDepth) return hipErrorInvalidValue;
    hipLaunchKernelGGL(KResetWhitted, dim3(1), dim3(1), 0, s, counter, totals);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    constexpr int B = 256;
    // lanes per pixel: enough pixel-lanes to fill the persistent grid (8 workgroups of 256
    // lanes per CU), at most 8; a large image keeps one lane per pixel (the group's lanes
    // repeat the pixel's closest-hit traces)
    const int64_t pixels = int64_t(a.W) * a.H;
    int G = 1;
    while (G < kWhittedLanesMax && pixels * G < int64_t(cu_count) * 8 * B) G *= 2;
    const int grid = TraceGrid(cu_count, pixels * G);
#define DPL_WHITTED(G_)                                                                                           \
    do {                                                                                                          \
        if (tree.ndirs > 0) {                                                                                     \
            if (totals) hipLaunchKernelGGL((KWhitted<true, B, true, G_>), dim3(grid), dim3(B), 0, s, a, counter, totals); \
            else hipLaunchKernelGGL((KWhitted<true, B, false, G_>), dim3(grid), dim3(B), 0, s, a, counter, totals);       \
        } else {                                                                                                  \
            if (totals) hipLaunchKernelGGL((KWhitted<false, B, true, G_>), dim3(grid), dim3(B), 0, s, a, counter, totals); \
            else hipLaunchKernelGGL((KWhitted<false, B, false, G_>), dim3(grid), dim3(B), 0, s, a, counter, totals);       \
        }                                                                                                         \
    } while (0)
    switch (G) {
        case 1: DPL_WHITTED(1); break;
        case 2: DPL_WHITTED(2); break;
        case 4: DPL_WHITTED(4); break;
        default: DPL_WHITTED(8); break;
    }
#undef DPL_WHITTED
    return hipGetLastError();
}

}  // namespace dpl
