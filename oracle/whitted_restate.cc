// Oracle — TEST INFRASTRUCTURE ONLY (see orc.h).
// Whitted ray tracer of src/renderer/cpu_raytracer.cc:50-288: primary rays from the
// reference's camera basis, area-light next-event estimation with hash-sampled light points
// (RandomAreaLightSample01 with seed 42, :28-39,181-182), shadow rays offset along the
// geometric normal (:41-46,214-229), Phong (:231-255), mirror reflection (:257-262) and
// refraction with total internal reflection (:264-287), to max_depth; emissive hits return
// their emission (:120-121). Area lights are the emissive triangles in triangle order, with
// the triangle's area (src/io/scene/obj_scene_loader.cc:95-105); r_ior = 1 / ior (:118).
//
// Intersections are the watertight closest hit / any hit of the oracle's trace (tmin 1e-6,
// ties to the lower primitive id), not the reference's Moller-Trumbore test.
//
// WhittedPixel evaluates the recursion
//   I(ray) = background | emission | L_local + I(reflected) * ks + I(refracted) * kt
// as a sum over the ray tree of (product of ks / kt along the path) x (local term), depth
// first with the reflected subtree before the refracted one — the order the GPU kernel uses.
// WhittedPixelRecursive evaluates the reference's recursion as written (bottom-up sums); the
// two differ only in the rounding of the colour sums, which the tests bound.
//
// x^y in the Phong term is PowPos (own log2 / exp2, relative error below 4e-6), so that the
// GPU kernel computes bit-identical colours; std::pow differs from it in the last bits.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "orc.h"

namespace orc {

namespace {

inline float AsFloat(uint32_t b) {
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}
inline uint32_t AsBits(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}

}  // namespace

// x^y for x >= 0, y >= 0: exp2(y log2 x). log2 by exponent extraction and the atanh series of
// the mantissa m in [sqrt(1/2), sqrt(2)): log2 m = (2/ln 2)(t + t^3/3 + ... + t^9/9),
// t = (m-1)/(m+1); exp2 of the fraction f in [0,1) by the Taylor series of e^(f ln 2) to
// degree 9; 2^n by the exponent bits. Results below 2^-126 flush to 0.
float PowPos(float x, float y) {
    if (y == 0.0f) return 1.0f;
    if (!(x >= 1.17549435e-38f)) return 0.0f;  // zero, denormal (or NaN): no contribution
    const uint32_t bits = AsBits(x);
    int e = int(bits >> 23) - 127;
    float m = AsFloat((bits & 0x7FFFFFu) | 0x3F800000u);
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
    const float lg = float(e) + (p * t) * 2.88539008f;  // 2 / ln 2
    const float z = y * lg;
    if (z < -126.0f) return 0.0f;
    if (z >= 128.0f) return AsFloat(0x7F800000u);
    const float fl = std::floor(z);
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
    return q * AsFloat(uint32_t(n + 127) << 23);
}

// src/util/util.h (HashU32, U01FromU32) keyed as in cpu_raytracer.cc:28-39
float AreaLightSample01(uint32_t seed, uint32_t px, uint32_t py, uint32_t light_id, uint32_t sample_idx,
                        uint32_t dimension) {
    uint32_t key = 0;
    key ^= seed * 0x9e3779b1U;
    key ^= px * 0x165667b1U;
    key ^= py * 0xd3a2646cU;
    key ^= light_id * 0x85ebca6bU;
    key ^= sample_idx * 0xc2b2ae35U;
    key ^= dimension * 0x27d4eb2fU;
    return U01FromU32(HashU32(key));
}

namespace {

template <typename F>
void ParallelRows(int64_t n, int threads, F f) {
    if (threads <= 1 || n < 1024) {
        for (int64_t i = 0; i < n; i++) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++)
        th.emplace_back([=, &f] {
            for (int64_t i = n * t / threads; i < n * (t + 1) / threads; i++) f(i);
        });
    for (auto& x : th) x.join();
}

struct WRay {
    Vec3 o, d;
    float tmax;
};

struct WHit {
    Vec3 pos, normal, geom_normal;
    float t, eps;
    int32_t prim;
};

struct WScene {
    const FinalTree* tree;
    const float* tri;        // n x 9, original order
    const float* vn;         // n x 9 vertex normals or null
    const int32_t* mat_id;   // n
    const WMaterial* mats;
    std::vector<int32_t> lights;  // emissive triangles in triangle order
    WParams p;
    Vec3 cam_pos;
};

inline Vec3 V(const float* a) { return Vec3{a[0], a[1], a[2]}; }
inline bool NonZero(const float* a) { return a[0] != 0.0f || a[1] != 0.0f || a[2] != 0.0f; }

Vec3 GeomNormal(const float* t) {  // triangle.h:22-26
    const Vec3 v0 = V(t), v1 = V(t + 3), v2 = V(t + 6);
    return Normalize(Cross(v1 - v0, v2 - v0));
}

float TriangleArea(const float* t) {  // triangle.h:28-34 (0.5 |e1 x e2|)
    const Vec3 v0 = V(t), v1 = V(t + 3), v2 = V(t + 6);
    return 0.5f * Length(Cross(v1 - v0, v2 - v0));
}

// A closest-hit query (AccelerationStructure::Intersect); st counts it and its traversal.
bool Closest(const WScene& S, const WRay& r, bool cull, WHit& h, WStats& st) {
    RayRec rr{r.o.x, r.o.y, r.o.z, kEpsilon, r.d.x, r.d.y, r.d.z, r.tmax};
    HitRec hr;
    TraceCount tc;
    TraceFinalOne(*S.tree, rr, cull ? 2u : 0u, hr, &tc);
    st.queries++, st.nodes += tc.nodes, st.tris += tc.tris;
    if (hr.prim < 0) return false;
    const float* t = S.tri + 9 * int64_t(hr.prim);
    h.t = hr.t;
    h.prim = hr.prim;
    h.pos = r.o + r.d * hr.t;
    const float b0 = (1.0f - hr.b1) - hr.b2;
    h.eps = RayEpsilon(h.pos, h.t);
    h.geom_normal = GeomNormal(t);
    h.normal = h.geom_normal;
    if (S.vn) {  // a triangle whose three vertex normals are all zero has none
        const float* n = S.vn + 9 * int64_t(hr.prim);
        if (NonZero(n) || NonZero(n + 3) || NonZero(n + 6))
            h.normal = Normalize((V(n) * b0 + V(n + 3) * hr.b1) + V(n + 6) * hr.b2);
    }
    return true;
}

bool AnyHit(const WScene& S, const WRay& r, WStats& st) {
    RayRec rr{r.o.x, r.o.y, r.o.z, kEpsilon, r.d.x, r.d.y, r.d.z, r.tmax};
    HitRec hr;
    TraceCount tc;
    TraceFinalOne(*S.tree, rr, 1u, hr, &tc);
    st.queries++, st.nodes += tc.nodes, st.tris += tc.tris;
    return hr.prim >= 0;
}

Vec3 OffsetRayOrigin(const Vec3& pos, float eps, const Vec3& gn, const Vec3& dir) {
    const Vec3 offset = Dot(dir, gn) > 0.0f ? gn : -gn;
    return pos + offset * eps;
}

bool IsShadowed(const WScene& S, const WHit& h, const Vec3& lpos, WStats& st) {
    st.shadow++;  // counted on entry, as cpu_raytracer.cc:223 does
    const Vec3 to_light = lpos - h.pos;
    const float dist = Length(to_light);
    if (dist < kEpsilon) return false;
    const Vec3 dir = to_light / dist;
    const WRay sr{OffsetRayOrigin(h.pos, h.eps, h.geom_normal, dir), dir, dist * (1.0f - h.eps)};
    return AnyHit(S, sr, st);
}

Vec3 Phong(const WScene& S, const WHit& h, const WMaterial& m, const Vec3& lpos, const Vec3& lcol) {
    const Vec3 d_l = Normalize(lpos - h.pos);
    const Vec3 d_v = Normalize(S.cam_pos - h.pos);
    const Vec3 d_r = Normalize(h.normal * 2.0f * Dot(h.normal, d_l) - d_l);
    const Vec3 I_a = lcol * Vec3{0.0f, 0.0f, 0.0f};
    const Vec3 I_d = lcol * V(m.diffuse) * std::max(0.0f, Dot(h.normal, d_l));
    const Vec3 I_s = lcol * V(m.specular) * PowPos(std::max(0.0f, Dot(d_v, d_r)), m.shininess);
    return (I_a + I_d) + I_s;
}

Vec3 LocalAreaLights(const WScene& S, int x, int y, const WHit& h, const WMaterial& m, WStats& st) {
    Vec3 color{0.0f, 0.0f, 0.0f};
    const int cnt = S.p.samples;
    for (size_t al = 0; al < S.lights.size(); al++) {
        const int32_t lt = S.lights[al];
        const float* ltri = S.tri + 9 * int64_t(lt);
        const Vec3 lcolor = V(S.mats[S.mat_id[lt]].emission);
        if (lt == h.prim) return lcolor;
        const float area = TriangleArea(ltri);
        const Vec3 lgn = GeomNormal(ltri);
        for (int s = 0; s < cnt; s++) {
            const float r1 = AreaLightSample01(42u, uint32_t(x), uint32_t(y), uint32_t(al), uint32_t(s), 0u);
            const float r2 = AreaLightSample01(42u, uint32_t(x), uint32_t(y), uint32_t(al), uint32_t(s), 1u);
            const float u = r1 + r2 > 1.0f ? 1.0f - r1 : r1;
            const float v = r1 + r2 > 1.0f ? 1.0f - r2 : r2;
            const Vec3 a = V(ltri), b = V(ltri + 3), c = V(ltri + 6);
            const Vec3 lpos = (a + (b - a) * u) + (c - a) * v;
            if (IsShadowed(S, h, lpos, st)) continue;
            const Vec3 to_light = lpos - h.pos;
            const float d = Length(to_light);
            const Vec3 d_l = to_light / d;
            const float w = ((area * std::max(0.0f, Dot(lgn, -d_l))) * std::max(0.0f, Dot(h.normal, d_l))) /
                            ((float(cnt) * d * d) + kEpsilon);
            color = color + Phong(S, h, m, lpos, lcolor * w);
        }
    }
    return color;
}

WRay ReflectionRay(const WRay& r, const WHit& h) {
    const Vec3 d_v = -r.d;
    const Vec3 dir = -d_v + (2.0f * Dot(d_v, h.normal)) * h.normal;
    return WRay{OffsetRayOrigin(h.pos, h.eps, h.geom_normal, dir), dir, r.tmax};
}

WRay RefractionRay(const WRay& r, const WHit& h, float ior, float r_ior) {
    const Vec3 v = -r.d;
    Vec3 n = h.normal;
    float eta = ior;
    float cos_i = Dot(v, h.normal);
    if (cos_i < 0.0f) {
        eta = r_ior;
        cos_i = -cos_i;
        n = -n;
    }
    const float sin2_i = std::max(0.0f, 1.0f - (cos_i * cos_i));
    const float sin2_t = sin2_i / (eta * eta);
    if (sin2_t >= 1.0f) return ReflectionRay(r, h);
    const float cos_t = std::sqrt(1.0f - sin2_t);
    const Vec3 dir = -v / eta + (cos_i / eta - cos_t) * n;
    return WRay{OffsetRayOrigin(h.pos, h.eps, h.geom_normal, dir), dir, r.tmax};
}

WRay PrimaryRay(const WScene& S, const float* basis, int x, int y) {  // cpu_raytracer.cc:99-106
    const Vec3 p00{basis[3], basis[4], basis[5]}, qw{basis[6], basis[7], basis[8]}, qh{basis[9], basis[10], basis[11]};
    const Vec3 d = Normalize((p00 + qw * float(x)) + qh * float(y));
    return WRay{S.cam_pos, d, S.p.cam[10]};
}

// Depth-first over the ray tree: reflection subtree before refraction subtree.
Vec3 WhittedPixel(const WScene& S, const float* basis, int x, int y, WStats& st) {
    struct Item {
        WRay r;
        Vec3 w;
        int depth;
    };
    Item stack[2 * kWhittedMaxDepth + 2];
    int sp = 0;
    stack[sp++] = Item{PrimaryRay(S, basis, x, y), Vec3{1.0f, 1.0f, 1.0f}, 0};
    Vec3 acc{0.0f, 0.0f, 0.0f};
    while (sp > 0) {
        const Item it = stack[--sp];
        WHit h;
        if (!Closest(S, it.r, S.p.cull != 0, h, st)) {
            acc = acc + it.w * V(S.p.background);
            continue;
        }
        const WMaterial& m = S.mats[S.mat_id[h.prim]];
        if (NonZero(m.emission)) {
            acc = acc + it.w * V(m.emission);
            continue;
        }
        acc = acc + it.w * LocalAreaLights(S, x, y, h, m, st);
        if (it.depth < S.p.max_depth) {
            if (NonZero(m.transmittance)) {
                stack[sp++] = Item{RefractionRay(it.r, h, m.ior, m.r_ior), it.w * V(m.transmittance), it.depth + 1};
                st.secondary++;
            }
            if (NonZero(m.specular)) {
                stack[sp++] = Item{ReflectionRay(it.r, h), it.w * V(m.specular), it.depth + 1};
                st.secondary++;
            }
        }
    }
    return acc;
}

// The reference's recursion as written (cpu_raytracer.cc:108-161).
Vec3 TraceRayRecursive(const WScene& S, int x, int y, const WRay& r, int depth, WStats& st) {
    WHit h;
    if (!Closest(S, r, S.p.cull != 0, h, st)) return V(S.p.background);
    const WMaterial& m = S.mats[S.mat_id[h.prim]];
    if (NonZero(m.emission)) return V(m.emission);
    Vec3 color = LocalAreaLights(S, x, y, h, m, st);
    if (depth < S.p.max_depth) {
        if (NonZero(m.specular)) {
            st.secondary++;
            color = color + TraceRayRecursive(S, x, y, ReflectionRay(r, h), depth + 1, st) * V(m.specular);
        }
        if (NonZero(m.transmittance)) {
            st.secondary++;
            color = color +
                    TraceRayRecursive(S, x, y, RefractionRay(r, h, m.ior, m.r_ior), depth + 1, st) * V(m.transmittance);
        }
    }
    return color;
}

}  // namespace

void WhittedRender(const FinalTree& tree, const float* tri, const float* vn, int64_t n, const int32_t* mat_id,
                   const WMaterial* mats, const WParams& p, bool recursive, float* image, int threads,
                   WStats* stats) {
    WScene S;
    S.tree = &tree;
    S.tri = tri;
    S.vn = vn;
    S.mat_id = mat_id;
    S.mats = mats;
    S.p = p;
    S.cam_pos = Vec3{p.cam[0], p.cam[1], p.cam[2]};
    for (int64_t i = 0; i < n; i++)
        if (NonZero(mats[mat_id[i]].emission)) S.lights.push_back(int32_t(i));
    float basis[12];
    CameraBasis(p.cam, p.W, p.H, basis);
    const int64_t pixels = int64_t(p.W) * p.H;
    std::vector<WStats> per(stats ? size_t(pixels) : 0);
    ParallelRows(pixels, threads, [&](int64_t i) {
        const int x = int(i % p.W), y = int(i / p.W);
        WStats st;
        const Vec3 c = recursive ? TraceRayRecursive(S, x, y, PrimaryRay(S, basis, x, y), 0, st)
                                 : WhittedPixel(S, basis, x, y, st);
        st.primary++;  // TracePixel (cpu_raytracer.cc:103)
        image[3 * i] = c.x, image[3 * i + 1] = c.y, image[3 * i + 2] = c.z;
        if (stats) per[size_t(i)] = st;
    });
    if (stats) {
        *stats = WStats{};
        for (const WStats& st : per) stats->Add(st);
    }
}

}  // namespace orc
