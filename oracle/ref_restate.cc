// Oracle — TEST INFRASTRUCTURE ONLY (see orc.h).
// Restatement of the reference's CPU geometry and SAH BVH, preserving operation order.
#include <algorithm>
#include <chrono>
#include <numeric>
#include <thread>
#include <utility>
#include <vector>

#include "orc.h"

namespace orc {

// src/scene/triangle.h:36 — (v0 + v1 + v2) / 3.0f, left-to-right, then a true division
Vec3 TriangleCentroid(const Tri& t) { return (t.v0 + t.v1 + t.v2) / 3.0f; }

// src/scene/triangle.h:38-44 — std::min({a,b,c}) keeps the first smallest element
AABB TriangleAabb(const Tri& t) {
    Vec3 mn{Min2(Min2(t.v0.x, t.v1.x), t.v2.x), Min2(Min2(t.v0.y, t.v1.y), t.v2.y),
            Min2(Min2(t.v0.z, t.v1.z), t.v2.z)};
    Vec3 mx{Max2(Max2(t.v0.x, t.v1.x), t.v2.x), Max2(Max2(t.v0.y, t.v1.y), t.v2.y),
            Max2(Max2(t.v0.z, t.v1.z), t.v2.z)};
    return AABB(mn, mx);
}

// src/scene/triangle.h:28-34 (the 0.5 literal is a double in the reference)
float TriangleArea(const Tri& t) {
    const Vec3 u = t.v1 - t.v0;
    const Vec3 v = t.v2 - t.v0;
    const Vec3 c = Cross(u, v);
    const float m = Length(c);
    return static_cast<float>(0.5 * m);
}

// src/scene/triangle.h:22-26
Vec3 TriangleGeomNormal(const Tri& t) { return Normalize(Cross(t.v1 - t.v0, t.v2 - t.v0)); }

// src/scene/triangle.cc:10-43 — Möller–Trumbore with an absolute determinant threshold
float IntersectRayMT(const Tri& tri, const Ray& ray, float& b1, float& b2, bool backface_culling, float eps) {
    const Vec3& a = tri.v0;
    const Vec3& b = tri.v1;
    const Vec3& c = tri.v2;
    const Vec3 e1(b - a), e2(c - a);
    const Vec3 pvec = Cross(ray.dir, e2);
    const float det = Dot(e1, pvec);
    if (backface_culling) {
        if (det < eps) return kInfinity;
    } else {
        if (std::fabs(det) < eps) return kInfinity;
    }
    const float inv_det = 1.0f / det;
    const Vec3 tvec = ray.origin - a;
    b1 = Dot(tvec, pvec) * inv_det;
    if (b1 < 0.0f || b1 > 1.0f) return kInfinity;
    const Vec3 qvec = Cross(tvec, e1);
    b2 = Dot(ray.dir, qvec) * inv_det;
    if (b2 < 0.0f || b1 + b2 > 1.0f) return kInfinity;
    return Dot(e2, qvec) * inv_det;
}

// src/scene/triangle.cc:45-53
Vec3 SampleSurface(const Tri& tri, float r1, float r2) {
    float u = r1 + r2 > 1.0f ? 1.0f - r1 : r1;
    float v = r1 + r2 > 1.0f ? 1.0f - r2 : r2;
    return tri.v0 + (tri.v1 - tri.v0) * u + (tri.v2 - tri.v0) * v;
}

// src/scene/ray.h:15-24
float RayEpsilon(const Vec3& p, float t) {
    float p_scale = std::max({std::fabs(p.x), std::fabs(p.y), std::fabs(p.z), 1.0f});
    float t_scale = std::max(t, 1.0f);
    float bias = std::max(1e-6f * p_scale, 1e-7f * t_scale);
    return std::clamp(bias, 1e-7f, 1e-3f);
}

// src/acceleration/aabb.h:40-55 — slab test, NaNs dropped by the std::max/min argument order
bool IntersectRayAabbRef(const Ray& ray, const AABB& b, float& t_min, float& t_max) {
    t_min = -kInfinity;
    t_max = kInfinity;
    for (int i = 0; i < 3; ++i) {
        float inv_d = 1.0f / ray.dir[i];
        float t0 = (b.min[i] - ray.origin[i]) * inv_d;
        float t1 = (b.max[i] - ray.origin[i]) * inv_d;
        if (inv_d < 0.0f) std::swap(t0, t1);
        t_min = std::max(t_min, t0);
        t_max = std::min(t_max, t1);
    }
    return t_max >= t_min;
}

// src/util/util.h:24-32
uint32_t HashU32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// src/util/util.h:34-37
float U01FromU32(uint32_t x) { return (x >> 8) * (1.0f / 16777216.0f); }

// ---------------------------------------------------------------------------------------
// SBvh — src/acceleration/sbvh.cc
// ---------------------------------------------------------------------------------------

// sbvh.cc:17-58
void SBvh::Build() {
    auto t0 = std::chrono::steady_clock::now();
    size_t n = tris.size();
    nodes.assign(2 * n - 1, SBvhNode{});
    tri_idx.resize(n);
    std::iota(tri_idx.begin(), tri_idx.end(), 0);

    std::vector<AABB> tbs;
    tbs.reserve(n);
    std::vector<Vec3> tcs;
    tcs.reserve(n);
    for (const Tri& t : tris) {
        tbs.emplace_back(TriangleAabb(t));
        tcs.emplace_back(TriangleCentroid(t));
    }
    AABB vb, cb;
    for (const AABB& tb : tbs) vb.Expand(tb);
    for (const Vec3& tc : tcs) cb.Expand(tc);

    SBvhNode& root = nodes[0];
    root.t_begin = 0;
    root.aabb_min = vb.min;
    root.aabb_max = vb.max;
    root.t_count = static_cast<int32_t>(n);
    next_node = 1;

    Subdivide(0, 0, cb, tcs, tbs);
    stats.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    CalculateStats();
}

// sbvh.cc:193-266 — 16-bin SAH on the longest centroid axis
float SBvh::FindBestSplit(SBvhNode& node, int& axis, float& split_pos, const AABB& cb, AABB& TB_L, AABB& TB_R,
                          const std::vector<Vec3>& tcs, const std::vector<AABB>& tbs) {
    axis = 0;
    if (cb.size.x >= cb.size.y && cb.size.x >= cb.size.z) {
        axis = 0;
    } else if (cb.size.y >= cb.size.z) {
        axis = 1;
    } else {
        axis = 2;
    }
    std::vector<AABB> bbs(bin_count, AABB());
    std::vector<size_t> ns(bin_count, 0);
    float k_0 = cb.min[axis];
    // int * double / float evaluated in double, then stored as float (sbvh.cc:213-214)
    float k_1 = static_cast<float>(bin_count * (1 - 1e-3) / cb.size[axis]);
    size_t t_end = node.t_begin + node.t_count;
    for (size_t t = node.t_begin; t < t_end; t++) {
        size_t ti = tri_idx[t];
        size_t bin = std::min(bin_count - 1, static_cast<int>(k_1 * (tcs[ti][axis] - k_0)));
        bbs[bin].Expand(tbs[ti]);
        ns[bin]++;
    }
    float best = kInfinity;
    int split_count = bin_count - 1;
    std::vector<size_t> N_Ls(split_count), N_Rs(split_count);
    std::vector<AABB> TB_Ls(split_count), TB_Rs(split_count);
    std::vector<float> A_Ls(split_count), A_Rs(split_count);
    N_Ls[0] = ns[0];
    TB_Ls[0] = bbs[0];
    A_Ls[0] = TB_Ls[0].SurfaceArea();
    for (int i = 1; i < split_count; i++) {
        N_Ls[i] = N_Ls[i - 1] + ns[i];
        TB_Ls[i] = TB_Ls[i - 1];
        TB_Ls[i].Expand(bbs[i]);
        A_Ls[i] = TB_Ls[i].SurfaceArea();
    }
    N_Rs[split_count - 1] = ns[split_count];
    TB_Rs[split_count - 1] = bbs[split_count];
    A_Rs[split_count - 1] = TB_Rs[split_count - 1].SurfaceArea();
    float scale = cb.size[axis] / bin_count;
    for (int i = split_count - 2; i >= 0; i--) {
        N_Rs[i] = N_Rs[i + 1] + ns[i + 1];
        TB_Rs[i] = TB_Rs[i + 1];
        TB_Rs[i].Expand(bbs[i + 1]);
        A_Rs[i] = TB_Rs[i].SurfaceArea();
        float cost = N_Ls[i] * A_Ls[i] + N_Rs[i] * A_Rs[i];
        if (cost != 0 && cost < best) {
            best = cost;
            split_pos = cb.min[axis] + scale * (i + 1);
            TB_L = TB_Ls[i];
            TB_R = TB_Rs[i];
        }
    }
    return best;
}

// sbvh.cc:268-341. With fix_child_bounds the child boxes are recomputed from the
// partitioned triangles instead of taken from the SAH bins (the bins and the partition
// plane disagree, so the faithful boxes need not enclose their triangles).
void SBvh::Subdivide(size_t node_idx, int depth, AABB cb, const std::vector<Vec3>& tcs, const std::vector<AABB>& tbs) {
    SBvhNode& node = nodes[node_idx];
    bool cb_too_small = cb.size.x < kEpsilon && cb.size.y < kEpsilon && cb.size.z < kEpsilon;
    if (node.t_count < max_leaf || depth >= max_depth || cb_too_small) return;

    int axis;
    float split_pos;
    AABB TB_L, TB_R;
    FindBestSplit(node, axis, split_pos, cb, TB_L, TB_R, tcs, tbs);

    int i = node.t_begin;
    int j = node.t_begin + node.t_count - 1;
    while (i <= j) {
        if (tcs[tri_idx[i]][axis] < split_pos) {
            i++;
        } else {
            std::swap(tri_idx[i], tri_idx[j]);
            j--;
        }
    }
    size_t N_L = i - node.t_begin;
    size_t N_R = node.t_count - N_L;
    if (N_L == 0 || N_R == 0) return;

    if (fix_child_bounds) {
        TB_L = AABB();
        TB_R = AABB();
        for (size_t k = node.t_begin; k < node.t_begin + N_L; k++) TB_L.Expand(tbs[tri_idx[k]]);
        for (size_t k = node.t_begin + N_L; k < size_t(node.t_begin + node.t_count); k++) TB_R.Expand(tbs[tri_idx[k]]);
    }

    size_t l = next_node++;
    nodes[l].t_begin = node.t_begin;
    nodes[l].t_count = static_cast<int32_t>(N_L);
    nodes[l].aabb_min = TB_L.min;
    nodes[l].aabb_max = TB_L.max;
    size_t r = next_node++;
    nodes[r].t_begin = static_cast<int32_t>(node.t_begin + N_L);
    nodes[r].t_count = static_cast<int32_t>(N_R);
    nodes[r].aabb_min = TB_R.min;
    nodes[r].aabb_max = TB_R.max;
    // `node` may not be used after the child writes above only if the vector reallocated;
    // nodes is sized 2n-1 up front, so the reference stays valid.
    node.t_begin = static_cast<int32_t>(l);
    node.t_count = 0;

    AABB CB_L, CB_R;
    for (size_t k = nodes[l].t_begin; k < size_t(nodes[l].t_begin + nodes[l].t_count); k++) CB_L.Expand(tcs[tri_idx[k]]);
    for (size_t k = nodes[r].t_begin; k < size_t(nodes[r].t_begin + nodes[r].t_count); k++) CB_R.Expand(tcs[tri_idx[k]]);
    Subdivide(l, depth + 1, CB_L, tcs, tbs);
    Subdivide(r, depth + 1, CB_R, tcs, tbs);
}

// sbvh.cc:343-367
void SBvh::CalculateStats() {
    stats.node_count = stats.inner_node_count = stats.leaf_node_count = 0;
    std::vector<size_t> s{0};
    while (!s.empty()) {
        size_t idx = s.back();
        s.pop_back();
        stats.node_count++;
        if (nodes[idx].is_leaf()) {
            stats.leaf_node_count++;
            continue;
        }
        stats.inner_node_count++;
        s.push_back(nodes[idx].t_begin);
        s.push_back(nodes[idx].t_begin + 1);
    }
    stats.memory_consumption = static_cast<int64_t>(nodes.size()) * 32;
}

static Ray ToRay(const RayRec& r) { return Ray{{r.ox, r.oy, r.oz}, {r.dx, r.dy, r.dz}, r.tmax}; }

// dummy_acceleration_structure.cc:21-56 (Intersect) and :58-79 (IntersectAny): every
// triangle in input order; closest keeps the first of equal t (strict <, from t_max), any-hit
// accepts t <= t_max. The reference's implicit t_min is kEpsilon.
static HitRec DummyOne(const std::vector<Tri>& tris, const RayRec& rr, bool any, bool cull) {
    const Ray ray = ToRay(rr);
    HitRec h{0.0f, 0.0f, 0.0f, -1};
    float t_hit = ray.t_max;
    for (size_t k = 0; k < tris.size(); k++) {
        float b1 = 0.0f, b2 = 0.0f;
        const float t = IntersectRayMT(tris[k], ray, b1, b2, cull);
        if (any) {
            if (t > kEpsilon && t <= ray.t_max) return HitRec{t, b1, b2, int32_t(k)};
            continue;
        }
        if (t > kEpsilon && t < t_hit) {
            t_hit = t;
            h = HitRec{t, b1, b2, int32_t(k)};
        }
    }
    return h;
}

void TraceDummy(const float* tri_xyz, int64_t ntri, const RayRec* rays, int64_t n, uint32_t flags, HitRec* hits,
                int threads) {
    std::vector<Tri> tris(ntri);
    for (int64_t k = 0; k < ntri; k++) tris[k] = LoadTri(tri_xyz + 9 * k);
    const bool any = flags & 1u, cull = flags & 2u;
    auto body = [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; i++) hits[i] = DummyOne(tris, rays[i], any, cull);
    };
    if (threads <= 1) {
        body(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++) th.emplace_back(body, n * t / threads, n * (t + 1) / threads);
    for (auto& x : th) x.join();
}

// sbvh.cc:60-145 (mode 0) and the engine's closest-hit rules over the same tree (mode 1)
bool SBvh::Intersect(const RayRec& rr, HitRec& h, bool cull, int mode, int64_t* visits) const {
    h = HitRec{0.0f, 0.0f, 0.0f, -1};
    const Ray ray = ToRay(rr);
    WtRay w = PrepareRay(rr);
    bool hit = false;
    float t_hit = mode == 0 ? ray.t_max : rr.tmax;
    int tri_hit = -1;
    float b1_hit = 0.0f, b2_hit = 0.0f;
    std::vector<size_t> stack;
    stack.reserve(64);
    stack.push_back(0);
    while (!stack.empty()) {
        size_t idx = stack.back();
        stack.pop_back();
        const SBvhNode& node = nodes[idx];
        if (visits) (*visits)++;
        if (node.is_leaf()) {
            for (int k = 0; k < node.t_count; k++) {
                int ti = tri_idx[node.t_begin + k];
                const Tri& tr = tris[ti];
                if (mode == 0) {
                    float b1, b2;
                    float t = IntersectRayMT(tr, ray, b1, b2, cull);
                    if (t > kEpsilon && t < t_hit) {
                        hit = true;
                        t_hit = t;
                        tri_hit = ti;
                        b1_hit = b1;
                        b2_hit = b2;
                    }
                } else {
                    float t, b1, b2;
                    const float v0[3] = {tr.v0.x, tr.v0.y, tr.v0.z};
                    const float v1[3] = {tr.v1.x, tr.v1.y, tr.v1.z};
                    const float v2[3] = {tr.v2.x, tr.v2.y, tr.v2.z};
                    if (IntersectWatertight(w, v0, v1, v2, cull, t, b1, b2) && t > rr.tmin &&
                        (t < t_hit || (t == t_hit && hit && ti < tri_hit))) {
                        hit = true;
                        t_hit = t;
                        tri_hit = ti;
                        b1_hit = b1;
                        b2_hit = b2;
                    }
                }
            }
            continue;
        }
        size_t li = node.t_begin, ri = li + 1;
        const SBvhNode& L = nodes[li];
        const SBvhNode& R = nodes[ri];
        bool lh, rh;
        float ltmin, ltmax, rtmin, rtmax;
        if (mode == 0) {
            lh = IntersectRayAabbRef(ray, AABB(L.aabb_min, L.aabb_max), ltmin, ltmax);
            rh = IntersectRayAabbRef(ray, AABB(R.aabb_min, R.aabb_max), rtmin, rtmax);
            if (ltmin >= t_hit || ltmax <= kEpsilon) lh = false;
            if (rtmin >= t_hit || rtmax <= kEpsilon) rh = false;
        } else {
            const float llo[3] = {L.aabb_min.x, L.aabb_min.y, L.aabb_min.z};
            const float lhi[3] = {L.aabb_max.x, L.aabb_max.y, L.aabb_max.z};
            const float rlo[3] = {R.aabb_min.x, R.aabb_min.y, R.aabb_min.z};
            const float rhi[3] = {R.aabb_max.x, R.aabb_max.y, R.aabb_max.z};
            lh = IntersectBoxRobust(w, rr.tmin, t_hit, llo, lhi, ltmin);
            rh = IntersectBoxRobust(w, rr.tmin, t_hit, rlo, rhi, rtmin);
        }
        if (lh && rh) {
            if (ltmin < rtmin) {
                stack.push_back(ri);
                stack.push_back(li);
            } else {
                stack.push_back(li);
                stack.push_back(ri);
            }
        } else if (lh) {
            stack.push_back(li);
        } else if (rh) {
            stack.push_back(ri);
        }
    }
    if (!hit) return false;
    h.t = t_hit;
    h.b1 = b1_hit;
    h.b2 = b2_hit;
    h.prim = tri_hit;
    return true;
}

// sbvh.cc:147-191 (mode 0: accept t < t_max strictly); mode 1 uses the engine's test
bool SBvh::IntersectAny(const RayRec& rr, bool cull, int mode) const {
    const Ray ray = ToRay(rr);
    WtRay w = PrepareRay(rr);
    std::vector<size_t> s{0};
    while (!s.empty()) {
        size_t idx = s.back();
        s.pop_back();
        const SBvhNode& node = nodes[idx];
        if (node.is_leaf()) {
            for (int k = 0; k < node.t_count; k++) {
                const Tri& tr = tris[tri_idx[node.t_begin + k]];
                if (mode == 0) {
                    float b1, b2;
                    float t = IntersectRayMT(tr, ray, b1, b2, cull);
                    if (t > kEpsilon && t < ray.t_max) return true;
                } else {
                    float t, b1, b2;
                    const float v0[3] = {tr.v0.x, tr.v0.y, tr.v0.z};
                    const float v1[3] = {tr.v1.x, tr.v1.y, tr.v1.z};
                    const float v2[3] = {tr.v2.x, tr.v2.y, tr.v2.z};
                    if (IntersectWatertight(w, v0, v1, v2, cull, t, b1, b2) && t > rr.tmin && t < rr.tmax) return true;
                }
            }
            continue;
        }
        size_t li = node.t_begin, ri = li + 1;
        for (size_t c : {li, ri}) {
            const SBvhNode& C = nodes[c];
            if (mode == 0) {
                float a, b;
                if (IntersectRayAabbRef(ray, AABB(C.aabb_min, C.aabb_max), a, b) && b > kEpsilon) s.push_back(c);
            } else {
                const float lo[3] = {C.aabb_min.x, C.aabb_min.y, C.aabb_min.z};
                const float hi[3] = {C.aabb_max.x, C.aabb_max.y, C.aabb_max.z};
                float tn;
                if (IntersectBoxRobust(w, rr.tmin, rr.tmax, lo, hi, tn)) s.push_back(c);
            }
        }
    }
    return false;
}

// SAH cost of the SBvh with the engine's formula (one formula for both trees).
double SahCostSBvh(const SBvh& b, double ct, double ci) {
    auto sa = [](const SBvhNode& nd) {
        double sx = double(nd.aabb_max.x) - nd.aabb_min.x, sy = double(nd.aabb_max.y) - nd.aabb_min.y,
               sz = double(nd.aabb_max.z) - nd.aabb_min.z;
        return 2.0 * (sx * sy + sy * sz + sx * sz);
    };
    double root = sa(b.nodes[0]);
    double sum = 0.0;
    std::vector<size_t> s{0};
    while (!s.empty()) {
        size_t idx = s.back();
        s.pop_back();
        const SBvhNode& nd = b.nodes[idx];
        if (nd.is_leaf()) {
            sum += ci * sa(nd) * nd.t_count;
        } else {
            sum += ct * sa(nd);
            s.push_back(nd.t_begin);
            s.push_back(nd.t_begin + 1);
        }
    }
    return root > 0 ? sum / root : 0.0;
}

}  // namespace orc
