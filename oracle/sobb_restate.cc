// Oracle — TEST INFRASTRUCTURE ONLY (see orc.h).
// SOBB node bounds (after Kacerik & Bittner 2025, "SOBB: Skewed Oriented Bounding Boxes for
// Ray Tracing"): a node is bounded by three slabs whose normals come from a fixed direction
// set (the 13 directions of the 26-DOP; kdop_size selects a subset). Every node keeps its
// k-DOP (min/max projection on every direction, padded by 1e-6 x the scene's largest
// coordinate so that FP32 rounding of the projections stays inside), unions are exact, and
// each child box of the final tree is the non-coplanar direction triple of least
// parallelepiped area  2 (w1 w2 + w2 w3 + w1 w3) / |det(n1, n2, n3)|  (ties: first triple).
// The reference has no implementation (README.md:154-161 lists SOBB as planned).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "orc.h"

namespace orc {

static const float kS2 = 0.70710677f;  // float(1/sqrt(2))
static const float kS3 = 0.57735026f;  // float(1/sqrt(3))
static const float kDirs13[13][3] = {
    {1, 0, 0},     {0, 1, 0},     {0, 0, 1},     {kS2, kS2, 0},  {kS2, -kS2, 0},  {kS2, 0, kS2},  {kS2, 0, -kS2},
    {0, kS2, kS2}, {0, kS2, -kS2}, {kS3, kS3, kS3}, {kS3, kS3, -kS3}, {kS3, -kS3, kS3}, {-kS3, kS3, kS3}};

SobbSet MakeSobbSet(int kdop_size) {
    SobbSet S;
    std::vector<int> pick;
    if (kdop_size >= 26) {
        for (int k = 0; k < 13; k++) pick.push_back(k);
    } else if (kdop_size >= 18) {
        for (int k = 0; k < 9; k++) pick.push_back(k);
    } else if (kdop_size >= 14) {
        pick = {0, 1, 2, 9, 10, 11, 12};
    } else {
        pick = {0, 1, 2};
    }
    S.ndirs = int(pick.size());
    for (int k = 0; k < S.ndirs; k++)
        for (int a = 0; a < 3; a++) S.dirs[k][a] = kDirs13[pick[k]][a];
    S.ntriples = 0;
    for (int a = 0; a < S.ndirs; a++)
        for (int b = a + 1; b < S.ndirs; b++)
            for (int c = b + 1; c < S.ndirs; c++) {
                const double n1[3] = {S.dirs[a][0], S.dirs[a][1], S.dirs[a][2]};
                const double n2[3] = {S.dirs[b][0], S.dirs[b][1], S.dirs[b][2]};
                const double n3[3] = {S.dirs[c][0], S.dirs[c][1], S.dirs[c][2]};
                const double cx = n2[1] * n3[2] - n2[2] * n3[1];
                const double cy = n2[2] * n3[0] - n2[0] * n3[2];
                const double cz = n2[0] * n3[1] - n2[1] * n3[0];
                const double det = n1[0] * cx + n1[1] * cy + n1[2] * cz;
                if (std::fabs(det) < 1e-3) continue;
                SobbSet::Triple& t = S.triples[S.ntriples++];
                t.a = a, t.b = b, t.c = c;
                t.inv_det = float(1.0 / std::fabs(det));
            }
    return S;
}

// n . v in FP32, left to right, no FMA
static inline float Proj(const float* n, const float* v) { return n[0] * v[0] + n[1] * v[1] + n[2] * v[2]; }

float KdopPad(const Bounds& b) {
    float s = 0.0f;
    for (int a = 0; a < 3; a++) s = Max2(s, Max2(std::fabs(b.vmin[a]), std::fabs(b.vmax[a])));
    // never zero, so that x - pad and x + pad are exact in sign and monotone in x
    return Max2(1e-6f * s, 1.40129846e-45f);
}

void LeafKdop(const SobbSet& S, const Tri& t, float pad, float* kd /*2*ndirs: lo..., hi...*/) {
    const float v0[3] = {t.v0.x, t.v0.y, t.v0.z}, v1[3] = {t.v1.x, t.v1.y, t.v1.z}, v2[3] = {t.v2.x, t.v2.y, t.v2.z};
    for (int k = 0; k < S.ndirs; k++) {
        const float p0 = Proj(S.dirs[k], v0), p1 = Proj(S.dirs[k], v1), p2 = Proj(S.dirs[k], v2);
        kd[k] = Min2(Min2(p0, p1), p2) - pad;
        kd[S.ndirs + k] = Max2(Max2(p0, p1), p2) + pad;
    }
}

int BestTriple(const SobbSet& S, const float* kd, float& best_sa) {
    float w[13];
    for (int k = 0; k < S.ndirs; k++) w[k] = kd[S.ndirs + k] - kd[k];
    int bt = 0;
    float best = 0.0f;
    for (int t = 0; t < S.ntriples; t++) {
        const SobbSet::Triple& T = S.triples[t];
        const float wa = w[T.a], wb = w[T.b], wc = w[T.c];
        const float sa = (wa * wb + wb * wc + wa * wc) * T.inv_det;
        if (t == 0 || sa < best) {
            best = sa;
            bt = t;
        }
    }
    best_sa = best;
    return bt;
}

// Final nodes with SOBB child bounds: (lo_a, hi_a, lo_b, hi_b, lo_c, hi_c) per child,
// child references as in the AABB layout, direction indices packed 4 bits each
// (left a,b,c in bits 0-11, right in bits 12-23).
void BuildSobbNodes(const SobbSet& S, const PlocResult& R, const std::vector<float>& kdop, int64_t n, int max_leaf,
                    std::vector<FinalTree::Node>& out) {
    // reuse the AABB layout's numbering: node i of R.tree.nodes corresponds to the same
    // PLOC node, so rebuild it with the same ordering rules
    const int64_t nn = 2 * n - 1;
    std::vector<uint8_t> inside(nn, 0);
    for (int64_t id = 0; id < nn; id++) {
        int q = R.parent[id];
        while (q >= 0 && R.count[q] <= max_leaf) {
            if (R.collapsed[q]) {
                inside[id] = 1;
                break;
            }
            q = R.parent[q];
        }
    }
    auto kept = [&](int64_t id) { return id >= n && !R.collapsed[id] && !inside[id]; };
    out = R.tree.nodes;  // same refs, boxes replaced below
    if (R.collapsed[nn - 1]) {
        // single-leaf tree: both slots bound the root
        FinalTree::Node& nd = out[0];
        float sa;
        const float* kd = &kdop[size_t(nn - 1) * 2 * S.ndirs];
        const int t = BestTriple(S, kd, sa);
        const SobbSet::Triple& T = S.triples[t];
        float* f = &nd.lx0;
        const int ids[3] = {T.a, T.b, T.c};
        for (int s = 0; s < 2; s++)
            for (int k = 0; k < 3; k++) {
                f[6 * s + 2 * k] = kd[ids[k]];
                f[6 * s + 2 * k + 1] = kd[S.ndirs + ids[k]];
            }
        const uint32_t d = uint32_t(T.a) | (uint32_t(T.b) << 4) | (uint32_t(T.c) << 8);
        nd.pad0 = int32_t(d | (d << 12));
        return;
    }
    int32_t K = 0;
    for (int64_t id = n; id < nn; id++) K += kept(id);
    int32_t rank = 0;
    for (int64_t id = n; id < nn; id++) {
        if (!kept(id)) continue;
        FinalTree::Node& nd = out[K - 1 - rank++];
        float* f = &nd.lx0;
        uint32_t dirs = 0;
        const int32_t ch[2] = {R.nodes[id].left, R.nodes[id].right};
        for (int s = 0; s < 2; s++) {
            const float* kd = &kdop[size_t(ch[s]) * 2 * S.ndirs];
            float sa;
            const SobbSet::Triple& T = S.triples[BestTriple(S, kd, sa)];
            const int ids[3] = {T.a, T.b, T.c};
            for (int k = 0; k < 3; k++) {
                f[6 * s + 2 * k] = kd[ids[k]];
                f[6 * s + 2 * k + 1] = kd[S.ndirs + ids[k]];
            }
            dirs |= (uint32_t(T.a) | (uint32_t(T.b) << 4) | (uint32_t(T.c) << 8)) << (12 * s);
        }
        nd.pad0 = int32_t(dirs);
    }
}

}  // namespace orc
