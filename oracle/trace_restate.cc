// Oracle — TEST INFRASTRUCTURE ONLY (see orc.h).
// Closest-hit / any-hit tracing with the watertight ray-triangle test of Woop, Benthin and
// Wald (JCGT 2013) and a conservative slab test (Ize 2013, "Robust BVH Ray Traversal").
// Closest-hit rule: the smallest t with tmin < t < tmax, ties to the lower primitive id —
// a result that does not depend on the tree or the traversal order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "orc.h"

namespace orc {

WtRay PrepareRay(const RayRec& r) {
    WtRay w;
    const float d[3] = {r.dx, r.dy, r.dz};
    float ax = std::fabs(r.dx), ay = std::fabs(r.dy), az = std::fabs(r.dz);
    int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (d[kz] < 0.0f) std::swap(kx, ky);
    w.kx = kx, w.ky = ky, w.kz = kz;
    w.Sx = d[kx] / d[kz];
    w.Sy = d[ky] / d[kz];
    w.Sz = 1.0f / d[kz];
    w.ox = r.ox, w.oy = r.oy, w.oz = r.oz;
    w.idx = 1.0f / r.dx;
    w.idy = 1.0f / r.dy;
    w.idz = 1.0f / r.dz;
    const float po = 4e-7f * (std::fabs(r.ox) + std::fabs(r.oy) + std::fabs(r.oz));
    const float o[3] = {r.ox, r.oy, r.oz}, id[3] = {w.idx, w.idy, w.idz};
    for (int a = 0; a < 3; a++) {
        const float p = id[a] >= 0.0f ? po : -po;
        w.bn[a] = o[a] + p;
        w.bf[a] = o[a] - p;
    }
    return w;
}

bool IntersectWatertight(const WtRay& w, const float* v0, const float* v1, const float* v2, bool cull, float& t,
                         float& b1, float& b2) {
    const float A[3] = {v0[0] - w.ox, v0[1] - w.oy, v0[2] - w.oz};
    const float B[3] = {v1[0] - w.ox, v1[1] - w.oy, v1[2] - w.oz};
    const float C[3] = {v2[0] - w.ox, v2[1] - w.oy, v2[2] - w.oz};
    const float Ax = A[w.kx] - w.Sx * A[w.kz];
    const float Ay = A[w.ky] - w.Sy * A[w.kz];
    const float Bx = B[w.kx] - w.Sx * B[w.kz];
    const float By = B[w.ky] - w.Sy * B[w.kz];
    const float Cx = C[w.kx] - w.Sx * C[w.kz];
    const float Cy = C[w.ky] - w.Sy * C[w.kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double CxBy = double(Cx) * double(By), CyBx = double(Cy) * double(Bx);
        U = float(CxBy - CyBx);
        double AxCy = double(Ax) * double(Cy), AyCx = double(Ay) * double(Cx);
        V = float(AxCy - AyCx);
        double BxAy = double(Bx) * double(Ay), ByAx = double(By) * double(Ax);
        W = float(BxAy - ByAx);
    }
    if (cull) {
        if (U < 0.0f || V < 0.0f || W < 0.0f) return false;
    } else {
        if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    }
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = w.Sz * A[w.kz];
    const float Bz = w.Sz * B[w.kz];
    const float Cz = w.Sz * C[w.kz];
    const float T = U * Az + V * Bz + W * Cz;
    if (det < 0.0f ? (T > 0.0f) : (T < 0.0f)) return false;
    t = T / det;
    b1 = V / det;
    b2 = W / det;
    return true;
}

// Near / far plane per axis by the sign of 1/d, both moved outward by po = 4e-7 |o|_1 (through
// the shifted origins bn / bf); a NaN distance (origin on a plane, d = +-0 on that axis:
// 0 * inf) is dropped by std::fmax / std::fmin, so the plane does not bound the ray (the
// reference's slab test drops it by its std::max / std::min argument order, aabb.h:40-55).
bool IntersectBoxRobust(const WtRay& w, float tmin, float tmax, const float lo[3], const float hi[3], float& tnear) {
    const float id[3] = {w.idx, w.idy, w.idz};
    float tn_[3], tf_[3];
    for (int a = 0; a < 3; a++) {
        const bool pos = id[a] >= 0.0f;
        tn_[a] = ((pos ? lo[a] : hi[a]) - w.bn[a]) * id[a];
        tf_[a] = ((pos ? hi[a] : lo[a]) - w.bf[a]) * id[a];
    }
    const float tn = std::fmax(std::fmax(tn_[0], tn_[1]), std::fmax(tn_[2], tmin));
    const float tf = std::fmin(std::fmin(tf_[0], tf_[1]), std::fmin(tf_[2], tmax));
    tnear = tn;
    return tn <= tf * 1.0000004f;
}

// SOBB slab test: per ray, n_k . o and 1 / (n_k . d) for every direction of the set, and
// an absolute pad po = 4e-7 |o|_1 that covers the rounding of n_k . o.
struct SobbRay {
    float on[13], idn[13];
    float po;
};

static SobbRay PrepareSobb(const SobbSet& S, const RayRec& r) {
    SobbRay s;
    for (int k = 0; k < S.ndirs; k++) {
        const float* n = S.dirs[k];
        s.on[k] = n[0] * r.ox + n[1] * r.oy + n[2] * r.oz;
        s.idn[k] = 1.0f / (n[0] * r.dx + n[1] * r.dy + n[2] * r.dz);
    }
    s.po = 4e-7f * (std::fabs(r.ox) + std::fabs(r.oy) + std::fabs(r.oz));
    return s;
}

static bool IntersectSobb(const SobbRay& s, float tmin, float tmax, const float* f, uint32_t dirs, float& tnear) {
    const int a = dirs & 15, b = (dirs >> 4) & 15, c = (dirs >> 8) & 15;
    const int k[3] = {a, b, c};
    float tn_[3], tf_[3];
    for (int q = 0; q < 3; q++) {  // near / far slab plane by the sign of 1 / (n . d), NaNs dropped
        const float lo = f[2 * q] - s.po, hi = f[2 * q + 1] + s.po, id = s.idn[k[q]], on = s.on[k[q]];
        const bool pos = id >= 0.0f;
        tn_[q] = ((pos ? lo : hi) - on) * id;
        tf_[q] = ((pos ? hi : lo) - on) * id;
    }
    const float tn = std::fmax(std::fmax(tn_[0], tn_[1]), std::fmax(tn_[2], tmin));
    const float tf = std::fmin(std::fmin(tf_[0], tf_[1]), std::fmin(tf_[2], tmax));
    tnear = tn;
    return tn <= tf * 1.0000004f;
}

static inline bool Better(float t, int32_t prim, float best_t, int32_t best_prim) {
    return t < best_t || (t == best_t && best_prim >= 0 && prim < best_prim);
}

void TraceFinalOne(const FinalTree& T, const RayRec& r, uint32_t flags, HitRec& h, TraceCount* count) {
    const bool any = flags & 1u, cull = flags & 2u;
    WtRay w = PrepareRay(r);
    SobbRay sr{};
    if (T.sobb) sr = PrepareSobb(T.set, r);
    h = HitRec{0.0f, 0.0f, 0.0f, -1};
    float best_t = r.tmax;
    int32_t stack[256];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        int32_t ref = stack[--sp];
        if (ref >= 0) {
            if (count) count->nodes++;
            const FinalTree::Node& nd = T.nodes[ref];
            float tl, tr;
            bool hl, hr;
            if (T.sobb) {
                const uint32_t dirs = uint32_t(nd.pad0);
                hl = IntersectSobb(sr, r.tmin, best_t, &nd.lx0, dirs & 0xFFFu, tl);
                hr = IntersectSobb(sr, r.tmin, best_t, &nd.lx0 + 6, (dirs >> 12) & 0xFFFu, tr);
            } else {
                const float llo[3] = {nd.lx0, nd.ly0, nd.lz0}, lhi[3] = {nd.lx1, nd.ly1, nd.lz1};
                const float rlo[3] = {nd.rx0, nd.ry0, nd.rz0}, rhi[3] = {nd.rx1, nd.ry1, nd.rz1};
                hl = IntersectBoxRobust(w, r.tmin, best_t, llo, lhi, tl);
                hr = IntersectBoxRobust(w, r.tmin, best_t, rlo, rhi, tr);
            }
            if (hl && hr) {
                if (tr < tl) {
                    stack[sp++] = nd.left;
                    stack[sp++] = nd.right;
                } else {
                    stack[sp++] = nd.right;
                    stack[sp++] = nd.left;
                }
            } else if (hl) {
                stack[sp++] = nd.left;
            } else if (hr) {
                stack[sp++] = nd.right;
            }
            if (sp > 250) std::abort();
        } else {
            int32_t packed = ~ref;
            int32_t begin = packed >> 4, cnt = packed & 15;
            for (int32_t k = begin; k < begin + cnt; k++) {
                const float* d = &T.tri_ordered[12 * size_t(k)];
                int32_t prim;
                std::memcpy(&prim, &d[3], 4);
                float t, b1, b2;
                if (count) count->tris++;
                if (!IntersectWatertight(w, d, d + 4, d + 8, cull, t, b1, b2)) continue;
                if (!(t > r.tmin)) continue;
                if (any) {
                    if (t < r.tmax) {
                        h = HitRec{t, b1, b2, prim};
                        return;
                    }
                    continue;
                }
                if (Better(t, prim, best_t, h.prim)) {
                    best_t = t;
                    h = HitRec{t, b1, b2, prim};
                }
            }
        }
    }
}

template <typename F>
static void ParallelFor(int64_t n, int threads, F f) {
    if (threads <= 1 || n < 1024) {
        for (int64_t i = 0; i < n; i++) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++) {
        th.emplace_back([=, &f] {
            for (int64_t i = n * t / threads; i < n * (t + 1) / threads; i++) f(i);
        });
    }
    for (auto& x : th) x.join();
}

void TraceFinal(const FinalTree& T, const RayRec* rays, int64_t n, uint32_t flags, HitRec* hits, int threads) {
    ParallelFor(n, threads, [&](int64_t i) { TraceFinalOne(T, rays[i], flags, hits[i]); });
}

void TraceFinalCounted(const FinalTree& T, const RayRec* rays, int64_t n, uint32_t flags, HitRec* hits, int threads,
                       TraceCount& total) {
    std::vector<TraceCount> per(static_cast<size_t>(n));
    ParallelFor(n, threads, [&](int64_t i) { TraceFinalOne(T, rays[i], flags, hits[i], &per[size_t(i)]); });
    for (const TraceCount& c : per) total.nodes += c.nodes, total.tris += c.tris;
}

void TraceWideCounted(const WideTree& T, const RayRec* rays, int64_t n, uint32_t flags, HitRec* hits, int threads,
                      TraceCount& total) {
    std::vector<TraceCount> per(static_cast<size_t>(n));
    ParallelFor(n, threads, [&](int64_t i) { TraceWideOne(T, rays[i], flags, hits[i], &per[size_t(i)]); });
    for (const TraceCount& c : per) total.nodes += c.nodes, total.tris += c.tris;
}

void TraceBrute(const float* tri_xyz, int64_t ntri, const RayRec* rays, int64_t n, uint32_t flags, HitRec* hits,
                int threads) {
    const bool any = flags & 1u, cull = flags & 2u;
    std::vector<float> canon(9 * ntri);
    for (int64_t i = 0; i < 9 * ntri; i++) canon[i] = Canon(tri_xyz[i]);
    ParallelFor(n, threads, [&](int64_t i) {
        const RayRec& r = rays[i];
        WtRay w = PrepareRay(r);
        HitRec h{0.0f, 0.0f, 0.0f, -1};
        float best_t = r.tmax;
        for (int64_t k = 0; k < ntri; k++) {
            const float* v = &canon[9 * k];
            float t, b1, b2;
            if (!IntersectWatertight(w, v, v + 3, v + 6, cull, t, b1, b2)) continue;
            if (!(t > r.tmin)) continue;
            if (any) {
                if (t < r.tmax) {
                    h = HitRec{t, b1, b2, int32_t(k)};
                    break;
                }
                continue;
            }
            if (Better(t, int32_t(k), best_t, h.prim)) {
                best_t = t;
                h = HitRec{t, b1, b2, int32_t(k)};
            }
        }
        hits[i] = h;
    });
}

// Exact-as-possible distance to the triangle's plane along the ray, in FP64, with the
// barycentrics of that point (Cramer's rule in double).
double IntersectFp64(const float* v, const RayRec& r, double& b1, double& b2) {
    double o[3] = {r.ox, r.oy, r.oz}, d[3] = {r.dx, r.dy, r.dz};
    double a[3] = {v[0], v[1], v[2]}, b[3] = {v[3], v[4], v[5]}, c[3] = {v[6], v[7], v[8]};
    double e1[3], e2[3], s[3];
    for (int k = 0; k < 3; k++) {
        e1[k] = b[k] - a[k];
        e2[k] = c[k] - a[k];
        s[k] = o[k] - a[k];
    }
    auto cross = [](const double* x, const double* y, double* z) {
        z[0] = x[1] * y[2] - x[2] * y[1];
        z[1] = x[2] * y[0] - x[0] * y[2];
        z[2] = x[0] * y[1] - x[1] * y[0];
    };
    auto dot = [](const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    double p[3], q[3];
    cross(d, e2, p);
    double det = dot(e1, p);
    if (det == 0.0) {
        b1 = b2 = 0.0;
        return HUGE_VAL;
    }
    cross(s, e1, q);
    b1 = dot(s, p) / det;
    b2 = dot(d, q) / det;
    return dot(e2, q) / det;
}

}  // namespace orc
