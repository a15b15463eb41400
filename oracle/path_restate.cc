// Oracle — TEST INFRASTRUCTURE ONLY (see orc.h).
// Diffuse path tracer (cosine-sampled bounces, PCG32 streams per pixel sample, emission
// collected on every hit) and its adjoint by path replay (Vicini, Speierer, Jakob 2021):
// the replay re-traces the path to recover its vertex materials and accumulates
//   dI/dE_m   += beta_k                 for every vertex k on material m
//   dI/drho_m += beta_k * S_k           with S_k = sum_{j>k} prod_{k<i<j} rho_{m_i} E_{m_j}
// using the backward recurrence S_{k-1} = E_{m_k} + rho_{m_k} * S_k.
// The camera follows src/renderer/cpu_raytracer.cc:62-73,99-106.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "orc.h"

namespace orc {

// PCG32 (O'Neill 2014, pcg32_srandom_r / pcg32_random_r)
uint32_t PcgNext(Pcg32& r) {
    uint64_t old = r.state;
    r.state = old * 6364136223846793005ull + r.inc;
    uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
void PcgSeed(Pcg32& r, uint64_t seed, uint64_t stream) {
    r.state = 0u;
    r.inc = (stream << 1u) | 1u;
    PcgNext(r);
    r.state += seed;
    PcgNext(r);
}
float PcgFloat(Pcg32& r) { return float(PcgNext(r) >> 8) * (1.0f / 16777216.0f); }

// sin and cos of 2*pi*u for u in [0,1): quadrant reduction, then Taylor polynomials of
// degree 11 (sin) and 12 (cos) on [0, pi/2) evaluated by Horner's rule.
void SinCos2Pi(float u, float& s, float& c) {
    float x = u * 4.0f;
    int q = int(x);
    if (q > 3) q = 3;
    float f = x - float(q);
    float th = f * 1.57079637f;
    float t2 = th * th;
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
    switch (q) {
        case 0: s = sp, c = cp; break;
        case 1: s = cp, c = -sp; break;
        case 2: s = -sp, c = -cp; break;
        default: s = -cp, c = sp; break;
    }
}

}  // namespace orc

namespace orc {

// Camera basis, src/renderer/cpu_raytracer.cc:62-73 (tan evaluated once on the host).
void CameraBasis(const float* cam /*pos3 dir3 up3 fov far*/, int w, int h, float* out12 /*pos p00 qw qh*/) {
    Vec3 pos{cam[0], cam[1], cam[2]}, dir{cam[3], cam[4], cam[5]}, up{cam[6], cam[7], cam[8]};
    const float fov = cam[9];
    const Vec3 right = Normalize(Cross(dir, up));
    const float gw = static_cast<float>(2.0 * std::tan(fov / 2.0f));
    const float gh = gw * (static_cast<float>(h) / static_cast<float>(w));
    const Vec3 qw = right * (gw / static_cast<float>(w - 1));
    const Vec3 qh = -up * (gh / static_cast<float>(h - 1));
    const Vec3 p00 = dir - right * (gw / 2) + up * (gh / 2);
    const Vec3 v[4] = {pos, p00, qw, qh};
    for (int k = 0; k < 4; k++) {
        out12[3 * k] = v[k].x;
        out12[3 * k + 1] = v[k].y;
        out12[3 * k + 2] = v[k].z;
    }
}

// One diffuse path: returns the number of vertices K (<= depth) and their materials and
// throughputs (product of the albedos of the earlier vertices). Every random number and
// every FP32 operation matches the HIP kernel.
int TracePathVertices(const FinalTree& T, const float* tri_xyz, const int32_t* mat_id, const float* albedo,
                      const float* basis, float far, int x, int y, int W, int spp, int s, int depth, uint64_t seed,
                      int32_t* mats, Vec3* betas) {
    Pcg32 rng;
    PcgSeed(rng, seed, (uint64_t(y) * uint64_t(W) + uint64_t(x)) * uint64_t(spp) + uint64_t(s));
    float fx = float(x), fy = float(y);
    if (spp > 1) {
        const float j1 = PcgFloat(rng);
        const float j2 = PcgFloat(rng);
        fx = fx + (j1 - 0.5f);
        fy = fy + (j2 - 0.5f);
    }
    const Vec3 pos{basis[0], basis[1], basis[2]}, p00{basis[3], basis[4], basis[5]};
    const Vec3 qw{basis[6], basis[7], basis[8]}, qh{basis[9], basis[10], basis[11]};
    Vec3 d = Normalize(p00 + (qw * fx) + (qh * fy));
    Vec3 o = pos;
    Vec3 beta(1.0f);
    int k = 0;
    for (int dep = 0; dep < depth; dep++) {
        RayRec r{o.x, o.y, o.z, kEpsilon, d.x, d.y, d.z, far};
        HitRec h;
        TraceFinal(T, &r, 1, 0u, &h, 1);
        if (h.prim < 0) break;
        const int32_t m = mat_id[h.prim];
        mats[k] = m;
        betas[k] = beta;
        k++;
        if (dep + 1 == depth) break;
        const Tri tr = LoadTri(tri_xyz + 9 * int64_t(h.prim));
        Vec3 n = Normalize(Cross(tr.v1 - tr.v0, tr.v2 - tr.v0));
        if (Dot(n, d) > 0.0f) n = -n;
        const Vec3 p = o + d * h.t;
        const float eps = RayEpsilon(p, h.t);
        o = p + n * eps;
        const float u1 = PcgFloat(rng);
        const float u2 = PcgFloat(rng);
        const float rr = std::sqrt(u1);
        float sn, cs;
        SinCos2Pi(u2, sn, cs);
        const float lx = rr * cs, ly = rr * sn, lz = std::sqrt(Max2(0.0f, 1.0f - u1));
        // orthonormal basis around n (Duff et al. 2017)
        const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
        const float a = -1.0f / (sign + n.z);
        const float b = n.x * n.y * a;
        const Vec3 t1{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
        const Vec3 t2{b, sign + n.y * n.y * a, -n.y};
        d = t1 * lx + t2 * ly + n * lz;
        beta = beta * Vec3{albedo[3 * m], albedo[3 * m + 1], albedo[3 * m + 2]};
    }
    return k;
}

}  // namespace orc
