// Oracle — TEST INFRASTRUCTURE ONLY (see orc.h). C entry points for ctypes.
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

#include "orc.h"

using namespace orc;

namespace {
struct PlocHandle {
    PlocResult r;
    int64_t n;
};
}  // namespace

extern "C" {

uint32_t orc_hash_u32(uint32_t x) { return HashU32(x); }
float orc_u01(uint32_t x) { return U01FromU32(x); }
float orc_ray_epsilon(float px, float py, float pz, float t) { return RayEpsilon(Vec3{px, py, pz}, t); }
void orc_centroid(const float* tri9, float* out3) {
    Tri t{{tri9[0], tri9[1], tri9[2]}, {tri9[3], tri9[4], tri9[5]}, {tri9[6], tri9[7], tri9[8]}};
    Vec3 c = TriangleCentroid(t);
    out3[0] = c.x, out3[1] = c.y, out3[2] = c.z;
}
float orc_triangle_area(const float* tri9) {
    Tri t{{tri9[0], tri9[1], tri9[2]}, {tri9[3], tri9[4], tri9[5]}, {tri9[6], tri9[7], tri9[8]}};
    return TriangleArea(t);
}

// Reference Möller–Trumbore (src/scene/triangle.cc:10-43); returns t (kInfinity on a miss).
float orc_intersect_mt(const float* tri9, const RayRec* r, int cull, float* b1, float* b2) {
    Tri t{{tri9[0], tri9[1], tri9[2]}, {tri9[3], tri9[4], tri9[5]}, {tri9[6], tri9[7], tri9[8]}};
    Ray ray{{r->ox, r->oy, r->oz}, {r->dx, r->dy, r->dz}, r->tmax};
    *b1 = *b2 = 0.0f;
    return IntersectRayMT(t, ray, *b1, *b2, cull != 0);
}
int orc_intersect_watertight(const float* tri9, const RayRec* r, int cull, float* t, float* b1, float* b2) {
    WtRay w = PrepareRay(*r);
    *t = *b1 = *b2 = 0.0f;
    return IntersectWatertight(w, tri9, tri9 + 3, tri9 + 6, cull != 0, *t, *b1, *b2) ? 1 : 0;
}
int orc_intersect_box(const RayRec* r, const float* lo, const float* hi, float tmax, float* tnear) {
    WtRay w = PrepareRay(*r);
    return IntersectBoxRobust(w, r->tmin, tmax, lo, hi, *tnear) ? 1 : 0;
}
int orc_intersect_box_ref(const RayRec* r, const float* lo, const float* hi, float* tmin, float* tmax) {
    Ray ray{{r->ox, r->oy, r->oz}, {r->dx, r->dy, r->dz}, r->tmax};
    return IntersectRayAabbRef(ray, AABB(Vec3{lo[0], lo[1], lo[2]}, Vec3{hi[0], hi[1], hi[2]}), *tmin, *tmax) ? 1 : 0;
}
double orc_intersect_fp64(const float* tri9, const RayRec* r, double* b1, double* b2) {
    return IntersectFp64(tri9, *r, *b1, *b2);
}

// ---- SBvh (reference restatement) ----
void* orc_sbvh_build(const float* tri, int64_t n, int max_depth, int max_leaf, int bin_count, int fix,
                     int32_t* stats3, double* build_ms) {
    if (n <= 0) return nullptr;
    SBvh* b = new SBvh();
    b->tris.resize(n);
    for (int64_t i = 0; i < n; i++) {
        const float* p = tri + 9 * i;
        b->tris[i] = Tri{{p[0], p[1], p[2]}, {p[3], p[4], p[5]}, {p[6], p[7], p[8]}};
    }
    b->max_depth = max_depth;
    b->max_leaf = max_leaf;
    b->bin_count = bin_count;
    b->fix_child_bounds = fix != 0;
    b->Build();
    if (stats3) {
        stats3[0] = b->stats.node_count;
        stats3[1] = b->stats.inner_node_count;
        stats3[2] = b->stats.leaf_node_count;
    }
    if (build_ms) *build_ms = b->stats.build_ms;
    return b;
}
void orc_sbvh_free(void* h) { delete static_cast<SBvh*>(h); }
int64_t orc_sbvh_node_capacity(void* h) { return int64_t(static_cast<SBvh*>(h)->nodes.size()); }
int64_t orc_sbvh_nodes_used(void* h) { return int64_t(static_cast<SBvh*>(h)->next_node); }
void orc_sbvh_export(void* h, float* minmax6, int32_t* begin_count2, int32_t* tri_idx) {
    SBvh* b = static_cast<SBvh*>(h);
    for (size_t i = 0; i < b->nodes.size(); i++) {
        const SBvhNode& nd = b->nodes[i];
        float* m = minmax6 + 6 * i;
        m[0] = nd.aabb_min.x, m[1] = nd.aabb_min.y, m[2] = nd.aabb_min.z;
        m[3] = nd.aabb_max.x, m[4] = nd.aabb_max.y, m[5] = nd.aabb_max.z;
        begin_count2[2 * i] = nd.t_begin;
        begin_count2[2 * i + 1] = nd.t_count;
    }
    std::memcpy(tri_idx, b->tri_idx.data(), b->tri_idx.size() * 4);
}
double orc_sbvh_sah(void* h) { return SahCostSBvh(*static_cast<SBvh*>(h)); }

// flags: bit0 any-hit, bit1 backface culling. mode 0 = reference rules, 1 = engine rules.
// Returns the wall-clock seconds spent tracing.
double orc_sbvh_trace(void* h, const RayRec* rays, int64_t n, uint32_t flags, int mode, HitRec* hits, int threads,
                      int64_t* visits) {
    SBvh* b = static_cast<SBvh*>(h);
    bool any = flags & 1u, cull = flags & 2u;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<int64_t> vis(threads > 0 ? threads : 1, 0);
    auto body = [&](int tid, int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; i++) {
            if (any) {
                bool hit = b->IntersectAny(rays[i], cull, mode);
                hits[i] = HitRec{0.0f, 0.0f, 0.0f, hit ? 0 : -1};
            } else {
                b->Intersect(rays[i], hits[i], cull, mode, &vis[tid]);
            }
        }
    };
    if (threads <= 1) {
        body(0, 0, n);
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < threads; t++) th.emplace_back(body, t, n * t / threads, n * (t + 1) / threads);
        for (auto& x : th) x.join();
    }
    double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (visits) {
        *visits = 0;
        for (auto v : vis) *visits += v;
    }
    return s;
}

// ---- engine build restatement ----
void orc_bounds(const float* tri, int64_t n, float* out12) {
    Bounds b = ComputeBounds(tri, n);
    std::memcpy(out12, b.vmin, 12);
    std::memcpy(out12 + 3, b.vmax, 12);
    std::memcpy(out12 + 6, b.cmin, 12);
    std::memcpy(out12 + 9, b.cmax, 12);
}
uint32_t orc_expand_bits10(uint32_t v) { return ExpandBits10(v); }
void orc_codes(const float* tri, int64_t n, int type, uint64_t* out) {
    Bounds b = ComputeBounds(tri, n);
    EmcLayout L = MakeEmcLayout(b);
    for (int64_t i = 0; i < n; i++) {
        Tri t = LoadTri(tri + 9 * i);
        Vec3 c = CentroidC(t);
        out[i] = type == 0 ? uint64_t(Morton30(c, b)) : Emc60(t, c, b, L);
    }
}
void orc_emc_layout(const float* bounds12, int32_t* nbits3, int8_t* axis_of48) {
    Bounds b;
    std::memcpy(b.vmin, bounds12, 12);
    std::memcpy(b.vmax, bounds12 + 3, 12);
    std::memcpy(b.cmin, bounds12 + 6, 12);
    std::memcpy(b.cmax, bounds12 + 9, 12);
    EmcLayout L = MakeEmcLayout(b);
    std::memcpy(nbits3, L.nbits, 12);
    std::memcpy(axis_of48, L.axis_of, 48);
}
void orc_ploc_nn(const float* lohi6, int32_t c, int32_t radius, int32_t* nn_out) {
    std::vector<PNode> cl(c);
    for (int i = 0; i < c; i++) {
        std::memcpy(cl[i].lo, lohi6 + 6 * i, 12);
        std::memcpy(cl[i].hi, lohi6 + 6 * i + 3, 12);
    }
    std::vector<int32_t> nn;
    PlocNearestNeighbours(cl, c, radius, nn);
    std::memcpy(nn_out, nn.data(), size_t(c) * 4);
}
float orc_sa_union(const float* alo, const float* ahi, const float* blo, const float* bhi) {
    return SurfaceAreaUnion(alo, ahi, blo, bhi);
}

void* orc_ploc_build(const float* tri, int64_t n, int morton_type, int radius, int max_leaf, int bv_type,
                     int kdop_size, double* build_ms) {
    if (n <= 0) return nullptr;
    PlocHandle* h = new PlocHandle();
    h->n = n;
    BuildParams p;
    p.morton_type = morton_type;
    p.radius = radius;
    p.max_leaf = max_leaf;
    p.bv_type = bv_type;
    p.kdop_size = kdop_size;
    auto t0 = std::chrono::steady_clock::now();
    PlocBuild(tri, n, p, h->r);
    if (build_ms) *build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return h;
}
void orc_ploc_free(void* h) { delete static_cast<PlocHandle*>(h); }
int32_t orc_ploc_iterations(void* h) { return int32_t(static_cast<PlocHandle*>(h)->r.iter_counts.size()); }
int64_t orc_ploc_final_nodes(void* h) { return int64_t(static_cast<PlocHandle*>(h)->r.tree.nodes.size()); }
// Any pointer may be null to skip that array. Sizes: codes n, idx n, nodes (2n-1)*8 words,
// parent 2n-1, iters orc_ploc_iterations, cost/count/collapsed 2n-1, fnodes 16 words each,
// prim_order n, tri_ordered 12n.
void orc_ploc_export(void* hv, uint64_t* codes, uint64_t* sorted_codes, uint32_t* sorted_idx, void* nodes,
                     int32_t* parent, int32_t* iters, float* cost, int32_t* count, uint8_t* collapsed, void* fnodes,
                     int32_t* prim_order, float* tri_ordered, float* bounds12) {
    PlocHandle* h = static_cast<PlocHandle*>(hv);
    const PlocResult& r = h->r;
    auto cp = [](void* dst, const void* src, size_t bytes) {
        if (dst && bytes) std::memcpy(dst, src, bytes);
    };
    cp(codes, r.codes.data(), r.codes.size() * 8);
    cp(sorted_codes, r.sorted_codes.data(), r.sorted_codes.size() * 8);
    cp(sorted_idx, r.sorted_idx.data(), r.sorted_idx.size() * 4);
    cp(nodes, r.nodes.data(), r.nodes.size() * sizeof(PNode));
    cp(parent, r.parent.data(), r.parent.size() * 4);
    cp(iters, r.iter_counts.data(), r.iter_counts.size() * 4);
    cp(cost, r.cost.data(), r.cost.size() * 4);
    cp(count, r.count.data(), r.count.size() * 4);
    cp(collapsed, r.collapsed.data(), r.collapsed.size());
    cp(fnodes, r.tree.nodes.data(), r.tree.nodes.size() * sizeof(FinalTree::Node));
    cp(prim_order, r.tree.prim_order.data(), r.tree.prim_order.size() * 4);
    cp(tri_ordered, r.tree.tri_ordered.data(), r.tree.tri_ordered.size() * 4);
    if (bounds12) {
        std::memcpy(bounds12, r.bounds.vmin, 12);
        std::memcpy(bounds12 + 3, r.bounds.vmax, 12);
        std::memcpy(bounds12 + 6, r.bounds.cmin, 12);
        std::memcpy(bounds12 + 9, r.bounds.cmax, 12);
    }
}
double orc_ploc_sah(void* h) { return SahCostFinal(static_cast<PlocHandle*>(h)->r.tree); }
// k-DOPs of every PLOC node (SOBB builds): (2n-1) x 2*ndirs floats; returns ndirs (0 if none).
int32_t orc_ploc_kdop(void* hv, float* out) {
    PlocHandle* h = static_cast<PlocHandle*>(hv);
    if (!h->r.tree.sobb) return 0;
    if (out) std::memcpy(out, h->r.kdop.data(), h->r.kdop.size() * 4);
    return h->r.tree.set.ndirs;
}
// Direction set: dirs (13x3), triples (286 x {a,b,c}), 1/|det| (286); returns the triple count.
int32_t orc_sobb_set(int32_t kdop_size, int32_t* ndirs, float* dirs, int32_t* abc, float* inv_det) {
    const SobbSet S = MakeSobbSet(kdop_size);
    *ndirs = S.ndirs;
    std::memcpy(dirs, S.dirs, sizeof(S.dirs));
    for (int t = 0; t < S.ntriples; t++) {
        abc[3 * t] = S.triples[t].a, abc[3 * t + 1] = S.triples[t].b, abc[3 * t + 2] = S.triples[t].c;
        inv_det[t] = S.triples[t].inv_det;
    }
    return S.ntriples;
}
void orc_ploc_trace(void* h, const RayRec* rays, int64_t n, uint32_t flags, HitRec* hits, int threads) {
    TraceFinal(static_cast<PlocHandle*>(h)->r.tree, rays, n, flags, hits, threads);
}
// The same trace with its traversal work: counts[0] inner nodes visited, counts[1] triangle tests.
void orc_ploc_trace_counted(void* h, const RayRec* rays, int64_t n, uint32_t flags, HitRec* hits, int threads,
                            int64_t* counts) {
    TraceCount tc;
    TraceFinalCounted(static_cast<PlocHandle*>(h)->r.tree, rays, n, flags, hits, threads, tc);
    counts[0] = tc.nodes, counts[1] = tc.tris;
}

// Trace over any tree in the final layout (e.g. one downloaded from the device).
// sobb_kdop: 0 for AABB child boxes, else the kdop_size of the SOBB direction set.
void orc_trace_final(const void* fnodes, int64_t nnodes, const float* tri_ordered, int64_t nprim, const RayRec* rays,
                     int64_t n, uint32_t flags, HitRec* hits, int threads, int32_t sobb_kdop) {
    FinalTree t;
    t.nodes.resize(nnodes);
    std::memcpy(t.nodes.data(), fnodes, size_t(nnodes) * sizeof(FinalTree::Node));
    t.tri_ordered.assign(tri_ordered, tri_ordered + 12 * nprim);
    if (sobb_kdop) {
        t.sobb = true;
        t.set = MakeSobbSet(sobb_kdop);
    }
    TraceFinal(t, rays, n, flags, hits, threads);
}
// orc_trace_final with the traversal work: counts[0] inner nodes visited, counts[1] triangle tests.
void orc_trace_final_counted(const void* fnodes, int64_t nnodes, const float* tri_ordered, int64_t nprim,
                             const RayRec* rays, int64_t n, uint32_t flags, HitRec* hits, int threads,
                             int32_t sobb_kdop, int64_t* counts) {
    FinalTree t;
    t.nodes.resize(nnodes);
    std::memcpy(t.nodes.data(), fnodes, size_t(nnodes) * sizeof(FinalTree::Node));
    t.tri_ordered.assign(tri_ordered, tri_ordered + 12 * nprim);
    if (sobb_kdop) {
        t.sobb = true;
        t.set = MakeSobbSet(sobb_kdop);
    }
    TraceCount tc;
    TraceFinalCounted(t, rays, n, flags, hits, threads, tc);
    counts[0] = tc.nodes, counts[1] = tc.tris;
}
// Axis-aligned child boxes of every node of a final tree (any bounding-volume type), from
// its triangles; out: nnodes x 64 B.
int orc_refit_aabb(const void* fnodes, int64_t nnodes, const float* tri_ordered, int64_t nprim, void* out) {
    FinalTree t;
    t.nodes.resize(nnodes);
    std::memcpy(t.nodes.data(), fnodes, size_t(nnodes) * sizeof(FinalTree::Node));
    t.tri_ordered.assign(tri_ordered, tri_ordered + 12 * nprim);
    std::vector<FinalTree::Node> o;
    try {
        RefitAabb(t, o);
    } catch (const std::exception&) {
        return -1;
    }
    std::memcpy(out, o.data(), o.size() * sizeof(FinalTree::Node));
    return 0;
}
// Wide nodes (width 4 or 8) from AABB final nodes; out: nnodes x orc_wide_node_bytes(width).
int orc_widen(const void* aabb_nodes, int64_t nnodes, int32_t width, void* out) {
    if (width != 4 && width != 8) return -1;
    std::vector<FinalTree::Node> nodes(static_cast<size_t>(nnodes));
    std::memcpy(nodes.data(), aabb_nodes, size_t(nnodes) * sizeof(FinalTree::Node));
    std::vector<uint8_t> o;
    Widen(nodes, width, o);
    std::memcpy(out, o.data(), o.size());
    return 0;
}
int32_t orc_wide_node_bytes(int32_t width) { return WideNodeBytes(width); }
// Trace over wide nodes; counts[0] wide nodes visited, counts[1] triangle tests (may be NULL).
void orc_trace_wide(const void* wnodes, int64_t nnodes, int32_t width, const float* tri_ordered, int64_t nprim,
                    const RayRec* rays, int64_t n, uint32_t flags, HitRec* hits, int threads, int64_t* counts) {
    WideTree t;
    t.width = width;
    t.node_count = nnodes;
    const uint8_t* b = static_cast<const uint8_t*>(wnodes);
    t.nodes.assign(b, b + size_t(nnodes) * size_t(WideNodeBytes(width)));
    t.tri_ordered.assign(tri_ordered, tri_ordered + 12 * nprim);
    TraceCount tc;
    TraceWideCounted(t, rays, n, flags, hits, threads, tc);
    if (counts) counts[0] = tc.nodes, counts[1] = tc.tris;
}
// SAH of a downloaded tree; root_box6 (lo xyz, hi xyz) is the tree's axis-aligned box.
double orc_sah_final(const void* fnodes, int64_t nnodes, const float* root_box6, int32_t sobb_kdop) {
    FinalTree t;
    t.nodes.resize(nnodes);
    std::memcpy(t.nodes.data(), fnodes, size_t(nnodes) * sizeof(FinalTree::Node));
    for (int a = 0; a < 3; a++) t.root_lo[a] = root_box6[a], t.root_hi[a] = root_box6[3 + a];
    if (sobb_kdop) {
        t.sobb = true;
        t.set = MakeSobbSet(sobb_kdop);
    }
    return SahCostFinal(t);
}
void orc_trace_brute(const float* tri, int64_t ntri, const RayRec* rays, int64_t n, uint32_t flags, HitRec* hits,
                     int threads) {
    TraceBrute(tri, ntri, rays, n, flags, hits, threads);
}
void orc_trace_dummy(const float* tri, int64_t ntri, const RayRec* rays, int64_t n, uint32_t flags, HitRec* hits,
                     int threads) {
    TraceDummy(tri, ntri, rays, n, flags, hits, threads);
}

// Whitted ray tracing over the PLOC tree of `ploc` (closest hits are tree-independent).
// mats: M x 16 floats (diffuse3 specular3 transmittance3 emission3 shininess ior r_ior pad).
// counts (may be NULL): primary, secondary, shadow, queries, inner nodes, triangle tests.
void orc_whitted(void* ploc, const float* tri_xyz, const float* vertex_normals, int64_t n, const int32_t* mat_id,
                 const float* mats, int W, int H, int max_depth, int samples, int cull, const float* background,
                 const float* cam, int recursive, float* image, int threads, int64_t* counts) {
    const FinalTree& T = static_cast<PlocHandle*>(ploc)->r.tree;
    WParams p{W, H, max_depth, samples, cull, {background[0], background[1], background[2]}, {}};
    for (int k = 0; k < 11; k++) p.cam[k] = cam[k];
    WStats st;
    WhittedRender(T, tri_xyz, vertex_normals, n, mat_id, reinterpret_cast<const WMaterial*>(mats), p, recursive != 0,
                  image, threads, counts ? &st : nullptr);
    if (counts) {
        const int64_t c[6] = {st.primary, st.secondary, st.shadow, st.queries, st.nodes, st.tris};
        std::memcpy(counts, c, sizeof(c));
    }
}
float orc_pow_pos(float x, float y) { return PowPos(x, y); }
float orc_area_light_sample01(uint32_t seed, uint32_t px, uint32_t py, uint32_t light_id, uint32_t s, uint32_t dim) {
    return AreaLightSample01(seed, px, py, light_id, s, dim);
}

// PCG32 / sincos KATs
void orc_pcg_stream(uint64_t seed, uint64_t stream, int32_t count, uint32_t* out) {
    Pcg32 r;
    PcgSeed(r, seed, stream);
    for (int i = 0; i < count; i++) out[i] = PcgNext(r);
}
void orc_sincos2pi(float u, float* s, float* c) { SinCos2Pi(u, *s, *c); }

}  // extern "C"

// ---- diffuse path tracing ----
extern "C" {

void orc_camera_basis(const float* cam, int w, int h, float* out12) { CameraBasis(cam, w, h, out12); }

// Forward render over the oracle's PLOC tree. image: W*H*3 (acc / spp per pixel, samples
// summed in order). path_mats (nullable): W*H*spp*depth vertex materials, -1 padded.
void orc_render(void* ploc, const float* tri_xyz, const int32_t* mat_id, const float* albedo, const float* emission,
                int W, int H, int spp, int depth, uint64_t seed, const float* cam, float* image, int32_t* path_mats,
                int threads) {
    const FinalTree& T = static_cast<PlocHandle*>(ploc)->r.tree;
    float basis[12];
    CameraBasis(cam, W, H, basis);
    const float far = cam[10];
    auto rows = [&](int y0, int y1) {
        std::vector<int32_t> mats(depth);
        std::vector<Vec3> betas(depth);
        for (int y = y0; y < y1; y++) {
            for (int x = 0; x < W; x++) {
                Vec3 acc(0.0f);
                for (int s = 0; s < spp; s++) {
                    int k = TracePathVertices(T, tri_xyz, mat_id, albedo, basis, far, x, y, W, spp, s, depth, seed,
                                              mats.data(), betas.data());
                    Vec3 L(0.0f);
                    for (int v = 0; v < k; v++) {
                        const int32_t m = mats[v];
                        L = L + betas[v] * Vec3{emission[3 * m], emission[3 * m + 1], emission[3 * m + 2]};
                    }
                    acc = acc + L;
                    if (path_mats) {
                        int32_t* pm = path_mats + ((int64_t(y) * W + x) * spp + s) * depth;
                        for (int v = 0; v < depth; v++) pm[v] = v < k ? mats[v] : -1;
                    }
                }
                float* px = image + (int64_t(y) * W + x) * 3;
                px[0] = acc.x / float(spp);
                px[1] = acc.y / float(spp);
                px[2] = acc.z / float(spp);
            }
        }
    };
    if (threads <= 1) {
        rows(0, H);
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < threads; t++) th.emplace_back(rows, H * t / threads, H * (t + 1) / threads);
        for (auto& x : th) x.join();
    }
}

// FP64 image of the recorded paths for parameters (albedo, emission) — the paths do not
// depend on them (albedo only scales throughput, it drives no sampling decision).
static void Image64(const int32_t* pm, int64_t pixels, int spp, int depth, const double* albedo,
                    const double* emission, double* img) {
    for (int64_t p = 0; p < pixels; p++) {
        double acc[3] = {0, 0, 0};
        for (int s = 0; s < spp; s++) {
            const int32_t* v = pm + (p * spp + s) * depth;
            double beta[3] = {1, 1, 1};
            for (int k = 0; k < depth && v[k] >= 0; k++) {
                for (int c = 0; c < 3; c++) {
                    acc[c] += beta[c] * emission[3 * v[k] + c];
                    beta[c] *= albedo[3 * v[k] + c];
                }
            }
        }
        for (int c = 0; c < 3; c++) img[3 * p + c] = acc[c] / spp;
    }
}

// L = sum (I - target)^2 in FP64.
double orc_path_loss_fp64(const int32_t* pm, int64_t pixels, int spp, int depth, const double* albedo,
                          const double* emission, const float* target) {
    std::vector<double> img(size_t(pixels) * 3);
    Image64(pm, pixels, spp, depth, albedo, emission, img.data());
    double l = 0.0;
    for (size_t i = 0; i < img.size(); i++) {
        const double r = img[i] - double(target[i]);
        l += r * r;
    }
    return l;
}

// Analytic dL/dalbedo and dL/demission in FP64 (M x 3 each, overwritten).
void orc_path_grad_fp64(const int32_t* pm, int64_t pixels, int spp, int depth, int M, const double* albedo,
                        const double* emission, const float* target, double* ga, double* ge) {
    std::vector<double> img(size_t(pixels) * 3);
    Image64(pm, pixels, spp, depth, albedo, emission, img.data());
    for (int i = 0; i < 3 * M; i++) ga[i] = ge[i] = 0.0;
    std::vector<double> beta(3 * size_t(depth));
    for (int64_t p = 0; p < pixels; p++) {
        double delta[3];
        for (int c = 0; c < 3; c++) delta[c] = 2.0 * (img[3 * p + c] - double(target[3 * p + c])) / spp;
        for (int s = 0; s < spp; s++) {
            const int32_t* v = pm + (p * spp + s) * depth;
            int K = 0;
            double b[3] = {1, 1, 1};
            while (K < depth && v[K] >= 0) {
                for (int c = 0; c < 3; c++) {
                    beta[3 * K + c] = b[c];
                    b[c] *= albedo[3 * v[K] + c];
                }
                K++;
            }
            double S[3] = {0, 0, 0};
            for (int k = K - 1; k >= 0; k--) {
                const int m = v[k];
                for (int c = 0; c < 3; c++) {
                    ge[3 * m + c] += delta[c] * beta[3 * k + c];
                    ga[3 * m + c] += delta[c] * beta[3 * k + c] * S[c];
                    S[c] = emission[3 * m + c] + albedo[3 * m + c] * S[c];
                }
            }
        }
    }
}

}  // extern "C"
