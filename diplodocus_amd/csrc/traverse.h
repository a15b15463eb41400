// Device-side traversal shared by the trace and render kernels: per-ray set-up for the
// watertight ray-triangle test (Woop, Benthin, Wald, JCGT 2013), the test itself, a
// conservative slab test (Ize 2013: the far distance is scaled by 1 + 2*gamma(3)), and an
// ordered depth-first traversal of the 64-byte two-child nodes with the stack in LDS.
//
// Closest-hit rule: smallest t with tmin < t < tmax, ties to the lower primitive id.
#pragma once

#include <hip/hip_runtime.h>

#include "dpl_internal.h"

namespace dpl {

// bn / bf: the origin shifted so that (plane - bn) and (plane - bf) are the distances to a
// box's near and far planes moved outward by po = 4e-7 |o|_1 (the pad of the SOBB slabs).
struct WtRay {
    int kx, ky, kz;
    float Sx, Sy, Sz;
    float ox, oy, oz;
    float idx, idy, idz;
    float tmin, tmax;
    float bnx, bny, bnz, bfx, bfy, bfz;
};

__device__ __forceinline__ float Sel3(float x, float y, float z, int k) { return k == 0 ? x : (k == 1 ? y : z); }

__device__ __forceinline__ WtRay PrepareRay(float ox, float oy, float oz, float tmin, float dx, float dy, float dz,
                                            float tmax) {
    WtRay w;
    const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    const int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    const float dkz = Sel3(dx, dy, dz, kz);
    if (dkz < 0.0f) {
        const int t = kx;
        kx = ky;
        ky = t;
    }
    w.kx = kx, w.ky = ky, w.kz = kz;
    w.Sx = Sel3(dx, dy, dz, kx) / dkz;
    w.Sy = Sel3(dx, dy, dz, ky) / dkz;
    w.Sz = 1.0f / dkz;
    w.ox = ox, w.oy = oy, w.oz = oz;
    w.idx = 1.0f / dx;
    w.idy = 1.0f / dy;
    w.idz = 1.0f / dz;
    w.tmin = tmin;
    w.tmax = tmax;
    const float po = 4e-7f * (fabsf(ox) + fabsf(oy) + fabsf(oz));
    const float px = w.idx >= 0.0f ? po : -po, py = w.idy >= 0.0f ? po : -po, pz = w.idz >= 0.0f ? po : -po;
    w.bnx = ox + px, w.bny = oy + py, w.bnz = oz + pz;
    w.bfx = ox - px, w.bfy = oy - py, w.bfz = oz - pz;
    return w;
}

__device__ __forceinline__ bool Watertight(const WtRay& w, const float4 p0, const float4 p1, const float4 p2,
                                           bool cull, float& t, float& b1, float& b2) {
    const float A0 = p0.x - w.ox, A1 = p0.y - w.oy, A2 = p0.z - w.oz;
    const float B0 = p1.x - w.ox, B1 = p1.y - w.oy, B2 = p1.z - w.oz;
    const float C0 = p2.x - w.ox, C1 = p2.y - w.oy, C2 = p2.z - w.oz;
    const float Akz = Sel3(A0, A1, A2, w.kz), Bkz = Sel3(B0, B1, B2, w.kz), Ckz = Sel3(C0, C1, C2, w.kz);
    const float Ax = Sel3(A0, A1, A2, w.kx) - w.Sx * Akz;
    const float Ay = Sel3(A0, A1, A2, w.ky) - w.Sy * Akz;
    const float Bx = Sel3(B0, B1, B2, w.kx) - w.Sx * Bkz;
    const float By = Sel3(B0, B1, B2, w.ky) - w.Sy * Bkz;
    const float Cx = Sel3(C0, C1, C2, w.kx) - w.Sx * Ckz;
    const float Cy = Sel3(C0, C1, C2, w.ky) - w.Sy * Ckz;
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

// Near and far planes are chosen per axis by the sign of 1/d (the same values as the min / max
// of the two plane distances, since both bounds move monotonically), each moved outward by the
// ray's pad po = 4e-7 |o|_1: the watertight test's t of a triangle the ray meets right at its
// origin while nearly parallel to it carries rounding that the box's exact planes do not, and
// without the pad the box is left before tmin where the triangle is hit after it. A ray
// parallel to an axis (d = +-0, 1/d = +-inf) whose origin lies exactly on a (padded) plane
// gives 0 * inf = NaN for that plane's distance; fmaxf / fminf drop a NaN operand, so that
// plane does not bound the ray, as for a ray running inside the slab (the reference's slab
// test drops such NaNs by its std::max / std::min argument order, src/acceleration/aabb.h:40-55).
// A min / max of the pair first would instead let the NaN discard the valid plane.
__device__ __forceinline__ bool Box(const WtRay& w, float tfar_limit, float x0, float x1, float y0, float y1, float z0,
                                    float z1, float& tnear) {
    const bool px = w.idx >= 0.0f, py = w.idy >= 0.0f, pz = w.idz >= 0.0f;
    const float tnx = ((px ? x0 : x1) - w.bnx) * w.idx, tfx = ((px ? x1 : x0) - w.bfx) * w.idx;
    const float tny = ((py ? y0 : y1) - w.bny) * w.idy, tfy = ((py ? y1 : y0) - w.bfy) * w.idy;
    const float tnz = ((pz ? z0 : z1) - w.bnz) * w.idz, tfz = ((pz ? z1 : z0) - w.bfz) * w.idz;
    const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, w.tmin));
    const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tfar_limit));
    tnear = tn;
    return tn <= tf * 1.0000004f;
}

// SOBB slab test. Per ray and direction k of the set: 1 / (n_k . d) in the lane's LDS column
// (idn[k * BLOCK]); n_k . o is recomputed from the workgroup's direction table (LDS, ndirs x
// 3) and the ray origin at each test, left to right without FMA (the oracle's order), which
// halves the per-lane LDS against keeping it. po = 4e-7 |o|_1 bounds the rounding of n_k . o
// and pads every slab.
struct SobbRay {
    const float* dirs;
    float* idn;
    float ox, oy, oz, po;
};

template <int BLOCK>
__device__ __forceinline__ void PrepareSobb(SobbRay& s, const float* dirs, int ndirs, float ox, float oy, float oz,
                                            float dx, float dy, float dz) {
    for (int k = 0; k < ndirs; k++) {
        const float nx = dirs[3 * k], ny = dirs[3 * k + 1], nz = dirs[3 * k + 2];
        s.idn[k * BLOCK] = 1.0f / (nx * dx + ny * dy + nz * dz);
    }
    s.dirs = dirs;
    s.ox = ox, s.oy = oy, s.oz = oz;
    s.po = 4e-7f * (fabsf(ox) + fabsf(oy) + fabsf(oz));
}

template <int BLOCK>
__device__ __forceinline__ bool SobbBox(const SobbRay& s, float tmin, float tfar_limit, float la, float ha,
                                        float lb, float hb, float lc, float hc, uint32_t dirs, float& tnear) {
    const int a = dirs & 15, b = (dirs >> 4) & 15, c = (dirs >> 8) & 15;
    auto on = [&](int k) { return s.dirs[3 * k] * s.ox + s.dirs[3 * k + 1] * s.oy + s.dirs[3 * k + 2] * s.oz; };
    const float oa = on(a), ia = s.idn[a * BLOCK];
    const float ob = on(b), ib = s.idn[b * BLOCK];
    const float oc = on(c), ic = s.idn[c * BLOCK];
    // near / far slab planes by the sign of 1 / (n . d), NaN distances dropped (as in Box)
    const float la_ = la - s.po, ha_ = ha + s.po, lb_ = lb - s.po, hb_ = hb + s.po, lc_ = lc - s.po, hc_ = hc + s.po;
    const bool pa = ia >= 0.0f, pb = ib >= 0.0f, pc = ic >= 0.0f;
    const float tna = ((pa ? la_ : ha_) - oa) * ia, tfa = ((pa ? ha_ : la_) - oa) * ia;
    const float tnb = ((pb ? lb_ : hb_) - ob) * ib, tfb = ((pb ? hb_ : lb_) - ob) * ib;
    const float tnc = ((pc ? lc_ : hc_) - oc) * ic, tfc = ((pc ? hc_ : lc_) - oc) * ic;
    const float tn = fmaxf(fmaxf(tna, tnb), fmaxf(tnc, tmin));
    const float tf = fminf(fminf(tfa, tfb), fminf(tfc, tfar_limit));
    tnear = tn;
    return tn <= tf * 1.0000004f;
}

struct HitResult {
    float t, b1, b2;
    int32_t prim;  // input triangle index, -1 on a miss
    int32_t slot;  // leaf-order position of the hit triangle
};

// Traversal stack: the top S entries live in the lane's LDS column (a ring of S slots,
// col[k * BLOCK]); deeper entries spill to a per-lane global region that is only touched
// when the stack is deeper than S (the spill region holds height - S + 1 entries).
template <int S, int BLOCK>
struct LaneStack {
    static constexpr int kBlock = BLOCK;
    int32_t* col;
    int32_t* spill;
    int sp;
    __device__ __forceinline__ void Push(int32_t x) {
        int32_t* slot = &col[(sp & (S - 1)) * BLOCK];
        if (sp >= S) spill[sp - S] = *slot;
        *slot = x;
        sp++;
    }
    __device__ __forceinline__ int32_t Pop() {
        sp--;
        int32_t* slot = &col[(sp & (S - 1)) * BLOCK];
        const int32_t x = *slot;
        if (sp >= S) *slot = spill[sp - S];
        return x;
    }
};

// The same ring with the spill region interleaved across lanes: entry d of lane l is
// spill[d * step + l] (step = lanes of the launch), so lanes of a wave that spill at the same
// depth write the same cache lines.
template <int S, int BLOCK>
struct LaneStackInterleaved {
    static constexpr int kBlock = BLOCK;
    int32_t* col;
    int32_t* spill;  // spill base + this lane's index
    int64_t step;
    int sp;
    __device__ __forceinline__ void Push(int32_t x) {
        int32_t* slot = &col[(sp & (S - 1)) * BLOCK];
        if (sp >= S) spill[int64_t(sp - S) * step] = *slot;
        *slot = x;
        sp++;
    }
    __device__ __forceinline__ int32_t Pop() {
        sp--;
        int32_t* slot = &col[(sp & (S - 1)) * BLOCK];
        const int32_t x = *slot;
        if (sp >= S) *slot = spill[int64_t(sp - S) * step];
        return x;
    }
};

struct TraceCounters {
    uint32_t nodes;  // inner nodes visited (two box tests each)
    uint32_t tris;   // triangle tests
};

// One triangle test of the closest-hit (or any-hit) rule; true when an any-hit query is done.
template <bool ANY>
__device__ __forceinline__ bool TestTriangle(const float4* tris, int32_t k, const WtRay& w, bool cull, float& best_t,
                                             HitResult& r) {
    const float4 p0 = tris[3 * int64_t(k)], p1 = tris[3 * int64_t(k) + 1], p2 = tris[3 * int64_t(k) + 2];
    float t, u1, u2;
    if (!Watertight(w, p0, p1, p2, cull, t, u1, u2)) return false;
    if (!(t > w.tmin)) return false;
    const int32_t prim = __float_as_int(p0.w);
    if (ANY) {
        if (t < w.tmax) {
            r = HitResult{t, u1, u2, prim, k};
            return true;
        }
        return false;
    }
    if (t < best_t || (t == best_t && r.prim >= 0 && prim < r.prim)) {
        best_t = t;
        r = HitResult{t, u1, u2, prim, k};
    }
    return false;
}

// Tests every triangle of the leaf reference `ref` (packed (first << 4) | count); true when
// an any-hit query is done.
template <bool ANY>
__device__ __forceinline__ bool TestLeaf(const float4* tris, int32_t prim_count, int32_t ref, const WtRay& w, bool cull,
                                         float& best_t, HitResult& r, TraceCounters& cnt) {
    const int32_t packed = ~ref;
    const int32_t b0 = packed >> 4;
    const int32_t np = b0 + (packed & 15) <= prim_count ? (packed & 15) : 0;
    for (int32_t k = b0; k < b0 + np; k++) {
        cnt.tris++;
        if (TestTriangle<ANY>(tris, k, w, cull, best_t, r)) return true;
    }
    return false;
}

// Ordered DFS: nearer child first, the other pushed. ANY stops at the first hit inside
// (tmin, tmax). Bounded: a tree visits each node at most once. SOBB selects the skewed
// child boxes (12 floats = left (lo hi) x 3 slabs, right the same; d.z packs the slab
// directions, 4 bits each, left in bits 0-11 and right in bits 12-23).
template <bool ANY, bool SOBB, class Stack, class SR>
__device__ __forceinline__ HitResult Traverse(const TraceTree& tree, const WtRay& w, const SR& sr, bool cull,
                                              Stack& st, TraceCounters& cnt) {
    HitResult r{0.0f, 0.0f, 0.0f, -1, -1};
    float best_t = w.tmax;
    st.sp = 0;
    int32_t ref = 0;
    int64_t budget = 2 * int64_t(tree.node_count) + 2;
    for (;;) {
        if (--budget < 0 || ref >= tree.node_count) break;
        if (ref >= 0) {
            cnt.nodes++;
            const float4* np = reinterpret_cast<const float4*>(tree.nodes + ref);
            const float4 a = np[0], b = np[1], c = np[2];
            const int4 d = reinterpret_cast<const int4*>(np)[3];
            float tl, tr;
            bool hl, hr;
            if (SOBB) {
                const uint32_t dirs = uint32_t(d.z);
                hl = SobbBox<Stack::kBlock>(sr, w.tmin, best_t, a.x, a.y, a.z, a.w, b.x, b.y, dirs & 0xFFFu, tl);
                hr = SobbBox<Stack::kBlock>(sr, w.tmin, best_t, b.z, b.w, c.x, c.y, c.z, c.w, (dirs >> 12) & 0xFFFu,
                                            tr);
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
                if (st.sp == 0) break;
                ref = st.Pop();
            }
        } else {
            if (TestLeaf<ANY>(tree.tris, tree.prim_count, ref, w, cull, best_t, r, cnt) || st.sp == 0) break;
            ref = st.Pop();
        }
    }
    return r;
}

}  // namespace dpl
