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
