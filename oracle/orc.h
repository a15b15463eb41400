// Oracle — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of (1) the reference's CPU acceleration/tracing path and (2) the
// algorithms of the GPU build-and-trace engine, used as the checker for the HIP kernels.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
// only as the checker or the timed CPU baseline, never as the product path.
//
// All stages that must be bit-exact against the GPU are written with explicit operation
// order and compiled with -ffp-contract=off (no FMA contraction) so that every FP32 result
// is the IEEE-754 round-to-nearest result of the same sequence of operations the HIP
// kernels perform.
//
// Parity status per function (pinning checks live in tests/):
//   Vec3/AABB/Triangle/MT/RayEpsilon/hash/camera  pinned by closed forms and hand KATs
//   SBvh build/intersect (faithful + fixed)       pinned by structural invariants and
//                                                 brute-force hit agreement
//   bounds, Morton30, EMC60                       pinned by hand KATs + invariants
//   stable radix-order                            pinned by std::stable_sort + brute force
//   PLOC, collapse, final layout                  pinned by invariants + brute-force NN on
//                                                 tiny inputs (no reference implementation
//                                                 exists: parity vs the reference unpinned)
//   watertight test, traversal                    pinned by FP64 intersection + brute force
//   wide trees (refit, widen, traversal)          pinned by exact refit of AABB trees,
//                                                 containment/tightness/greedy-rule checks
//                                                 and hit agreement with the two-child trace
//   path tracer forward/adjoint                   pinned by the furnace closed form and
//                                                 FP64 central differences
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------------------
// Reference value types (src/util/vec3.h:13-124, src/util/util.h:15-37)
// ---------------------------------------------------------------------------------------
constexpr float kEpsilon = 1e-6f;                               // util.h:15
constexpr float kInfinity = std::numeric_limits<float>::max();  // util.h:16

struct Vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    Vec3() = default;
    Vec3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
    explicit Vec3(float v) : x(v), y(v), z(v) {}
    float& operator[](int i) { return i == 0 ? x : i == 1 ? y : z; }
    const float& operator[](int i) const { return i == 0 ? x : i == 1 ? y : z; }
};
inline Vec3 operator+(Vec3 a, const Vec3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator-(Vec3 a, const Vec3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator*(Vec3 a, const Vec3& b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline Vec3 operator*(Vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline Vec3 operator*(float s, Vec3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline Vec3 operator/(Vec3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline Vec3 operator-(const Vec3& a) { return {-a.x, -a.y, -a.z}; }
inline bool operator==(const Vec3& a, const Vec3& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
inline bool operator!=(const Vec3& a, const Vec3& b) { return !(a == b); }
// vec3.h:118 — left-to-right sum of products, no FMA
inline float Dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// vec3.h:119-121
inline Vec3 Cross(const Vec3& a, const Vec3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float Length(const Vec3& v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }  // vec3.h:109
// vec3.h:111-117: zero vector when the length is at most eps, else divide by the length
inline Vec3 Normalize(Vec3 v, float eps = kEpsilon) {
    float len = Length(v);
    if (len <= eps) return Vec3(0.0f);
    return v / len;
}
inline Vec3 Abs(const Vec3& v) { return {std::fabs(v.x), std::fabs(v.y), std::fabs(v.z)}; }

// std::min / std::max semantics (return the first argument unless the second compares
// strictly smaller / larger). The GPU kernels use the same select form.
inline float Min2(float a, float b) { return (b < a) ? b : a; }
inline float Max2(float a, float b) { return (a < b) ? b : a; }

// src/acceleration/aabb.h:9-38. `size` of a default-constructed box is (0,0,0) and is only
// recomputed by Expand or the (min,max) constructor, exactly as in the reference.
struct AABB {
    Vec3 min{kInfinity, kInfinity, kInfinity};
    Vec3 max{-kInfinity, -kInfinity, -kInfinity};
    Vec3 size{0.0f, 0.0f, 0.0f};
    AABB() = default;
    AABB(const Vec3& mn, const Vec3& mx) : min(mn), max(mx), size(Abs(mx - mn)) {}
    float SurfaceArea() const { return 2.0f * (size.x * size.y + size.y * size.z + size.x * size.z); }
    void Expand(const AABB& o) {
        min.x = Min2(min.x, o.min.x);
        min.y = Min2(min.y, o.min.y);
        min.z = Min2(min.z, o.min.z);
        max.x = Max2(max.x, o.max.x);
        max.y = Max2(max.y, o.max.y);
        max.z = Max2(max.z, o.max.z);
        size = max - min;
    }
    void Expand(const Vec3& p) {
        min.x = Min2(min.x, p.x);
        min.y = Min2(min.y, p.y);
        min.z = Min2(min.z, p.z);
        max.x = Max2(max.x, p.x);
        max.y = Max2(max.y, p.y);
        max.z = Max2(max.z, p.z);
        size = max - min;
    }
};

// src/scene/ray.h:7-13; the reference's t_min is the implicit kEpsilon.
struct Ray {
    Vec3 origin, dir;
    float t_max;
};

// Position-only triangle (the build and trace need only positions: 36 B per triangle).
struct Tri {
    Vec3 v0, v1, v2;
};

// ---------------------------------------------------------------------------------------
// C-ABI-shaped records shared with the engine (32 B ray, 16 B hit).
// ---------------------------------------------------------------------------------------
struct RayRec {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
};
struct HitRec {
    float t, b1, b2;
    int32_t prim;  // -1 = miss
};

// ---------------------------------------------------------------------------------------
// Reference restatement: triangle helpers and Möller–Trumbore (src/scene/triangle.{h,cc})
// ---------------------------------------------------------------------------------------
Vec3 TriangleCentroid(const Tri& t);   // triangle.h:36
AABB TriangleAabb(const Tri& t);       // triangle.h:38-44
float TriangleArea(const Tri& t);      // triangle.h:28-34
Vec3 TriangleGeomNormal(const Tri& t); // triangle.h:22-26
float IntersectRayMT(const Tri& tri, const Ray& ray, float& b1, float& b2, bool backface_culling,
                     float eps = kEpsilon);  // triangle.cc:10-43
Vec3 SampleSurface(const Tri& tri, float r1, float r2);  // triangle.cc:45-53
float RayEpsilon(const Vec3& p, float t);                 // ray.h:15-24
bool IntersectRayAabbRef(const Ray& ray, const AABB& b, float& t_min, float& t_max);  // aabb.h:40-55
uint32_t HashU32(uint32_t x);                              // util.h:24-32
float U01FromU32(uint32_t x);                              // util.h:34-37

// ---------------------------------------------------------------------------------------
// Reference restatement: SBvh (src/acceleration/sbvh.{h,cc})
// ---------------------------------------------------------------------------------------
struct SBvhNode {  // sbvh.h:25-31
    Vec3 aabb_min{kInfinity, kInfinity, kInfinity};
    Vec3 aabb_max{-kInfinity, -kInfinity, -kInfinity};
    int32_t t_begin = 0;
    int32_t t_count = 0;
    bool is_leaf() const { return t_count > 0; }
};

struct SBvhStats {
    int node_count = 0, inner_node_count = 0, leaf_node_count = 0;
    int64_t memory_consumption = 0;
    double build_ms = 0.0;
};

struct SBvh {
    std::vector<Tri> tris;
    std::vector<int32_t> tri_idx;
    std::vector<SBvhNode> nodes;
    size_t next_node = 0;
    int max_depth = 32, max_leaf = 4, bin_count = 16;
    bool fix_child_bounds = false;
    SBvhStats stats;

    void Build();
    // mode 0: faithful reference rules (MT test, strict t<t_hit, ties to first visited)
    // mode 1: engine rules (watertight test, ties to the lower primitive id)
    bool Intersect(const RayRec& r, HitRec& h, bool backface_culling, int mode, int64_t* visits) const;
    bool IntersectAny(const RayRec& r, bool backface_culling, int mode) const;

   private:
    float FindBestSplit(SBvhNode& node, int& axis, float& split_pos, const AABB& cb, AABB& TB_L, AABB& TB_R,
                        const std::vector<Vec3>& tcs, const std::vector<AABB>& tbs);
    void Subdivide(size_t node_idx, int depth, AABB cb, const std::vector<Vec3>& tcs, const std::vector<AABB>& tbs);
    void CalculateStats();
};

// ---------------------------------------------------------------------------------------
// Engine algorithm restatement: codes, sort, PLOC, collapse, final layout
// ---------------------------------------------------------------------------------------
// Canonicalise -0.0 to +0.0 (x + 0.0f). Applied to every loaded vertex coordinate on both
// sides so that min/max and packed-distance comparisons never see a negative zero.
inline float Canon(float x) { return x + 0.0f; }
Tri LoadTri(const float* xyz9);

struct Bounds {
    float vmin[3], vmax[3];  // vertex (scene) bounds
    float cmin[3], cmax[3];  // centroid bounds
};
Bounds ComputeBounds(const float* tri_xyz, int64_t n);
Vec3 CentroidC(const Tri& t);  // ((v0+v1)+v2)/3.0f per component

uint32_t ExpandBits10(uint32_t v);
uint32_t Morton30(const Vec3& c, const Bounds& b);

struct EmcLayout {
    int nbits[3];         // spatial bits per axis
    int8_t axis_of[48];   // axis of the t-th spatial bit (MSB first)
    float scale[3];       // 2^nbits / extent (0 for a zero extent)
};
EmcLayout MakeEmcLayout(const Bounds& b);
uint64_t Emc60(const Tri& t, const Vec3& c, const Bounds& b, const EmcLayout& L);

struct BuildParams {
    int morton_type = 0;  // 0 = Morton30, 1 = EMC60
    int bv_type = 0;      // 0 = AABB, 1 = SOBB
    int radius = 16;
    int max_leaf = 4;
    int kdop_size = 32;
};

struct PNode {       // PLOC node (leaves 0..n-1 in sorted order, internal nodes n..2n-2)
    float lo[3], hi[3];
    int32_t left, right;  // for a leaf: left = original triangle index, right = -1
};

// SOBB direction set: unit normals and the non-coplanar direction triples with 1/|det|.
struct SobbSet {
    int ndirs = 0;
    float dirs[13][3];
    int ntriples = 0;
    struct Triple {
        int a, b, c;
        float inv_det;
    } triples[286];
};
SobbSet MakeSobbSet(int kdop_size);

struct FinalTree {
    // 64-byte two-child node: child boxes + child references. A reference r >= 0 is an
    // inner node index; r < 0 is a leaf whose ~r packs (prim_begin << 4) | prim_count.
    // AABB boxes: (lx0 lx1 ly0 ly1 | rx0 rx1 ry0 ry1 | lz0 lz1 rz0 rz1).
    // SOBB boxes: 12 floats = left (lo_a hi_a lo_b hi_b lo_c hi_c), right (same); pad0
    // packs the direction indices 4 bits each (left a,b,c bits 0-11, right bits 12-23).
    struct Node {
        float lx0, lx1, ly0, ly1;
        float rx0, rx1, ry0, ry1;
        float lz0, lz1, rz0, rz1;
        int32_t left, right, pad0, pad1;
    };
    std::vector<Node> nodes;
    std::vector<int32_t> prim_order;  // leaf-order position -> original triangle
    std::vector<float> tri_ordered;   // 12 floats per prim: v0 xyz,id | v1 xyz,0 | v2 xyz,0
    bool sobb = false;
    SobbSet set;
    float root_lo[3] = {0, 0, 0}, root_hi[3] = {0, 0, 0};  // axis-aligned box of the whole tree
};

struct PlocResult {
    Bounds bounds;
    std::vector<uint64_t> codes;        // per triangle, input order
    std::vector<uint64_t> sorted_codes;
    std::vector<uint32_t> sorted_idx;
    std::vector<PNode> nodes;           // 2n-1
    std::vector<int32_t> parent;        // 2n-1 (root: -1)
    std::vector<int32_t> iter_counts;   // cluster count at the start of every iteration
    std::vector<float> cost;            // collapse SAH cost per node
    std::vector<int32_t> count;         // triangles below each node
    std::vector<uint8_t> collapsed;     // node becomes a leaf
    std::vector<float> kdop;            // SOBB builds: 2 * ndirs floats per node (lo..., hi...)
    FinalTree tree;
};

float KdopPad(const Bounds& b);
void LeafKdop(const SobbSet& S, const Tri& t, float pad, float* kd);
int BestTriple(const SobbSet& S, const float* kd, float& best_sa);
void BuildSobbNodes(const SobbSet& S, const PlocResult& R, const std::vector<float>& kdop, int64_t n, int max_leaf,
                    std::vector<FinalTree::Node>& out);

void PlocBuild(const float* tri_xyz, int64_t n, const BuildParams& p, PlocResult& out);
void PlocNearestNeighbours(const std::vector<PNode>& clusters_box, int c, int radius, std::vector<int32_t>& nn);
float SurfaceAreaUnion(const float* alo, const float* ahi, const float* blo, const float* bhi);
uint32_t PairHash(int32_t i, int32_t j);

// SAH cost: Ct * sum_inner SA + Ci * sum_leaf SA * |leaf|, divided by SA(root), FP64 sum.
double SahCostFinal(const FinalTree& t, double ct = 1.0, double ci = 1.0);
double SahCostSBvh(const SBvh& b, double ct = 1.0, double ci = 1.0);

// ---------------------------------------------------------------------------------------
// Trace: watertight triangle test (Woop, Benthin, Wald 2013), robust slab test, traversal
// ---------------------------------------------------------------------------------------
struct WtRay {
    int kx, ky, kz;
    float Sx, Sy, Sz;
    float ox, oy, oz;
    float idx, idy, idz;
    // origins for the box's near / far plane distances, planes padded by po = 4e-7 |o|_1
    float bn[3], bf[3];
};
WtRay PrepareRay(const RayRec& r);
// Returns true and (t, b1, b2) when the ray hits the triangle; no t-range check.
bool IntersectWatertight(const WtRay& w, const float* v0, const float* v1, const float* v2, bool backface_culling,
                         float& t, float& b1, float& b2);
bool IntersectBoxRobust(const WtRay& w, float tmin, float tmax, const float lo[3], const float hi[3], float& tnear);

void TraceFinal(const FinalTree& t, const RayRec* rays, int64_t n, uint32_t flags, HitRec* hits, int threads);
// Traversal work of a trace (the counterparts of AccelerationStats::traversal_count and
// intersection_count, src/stats/acceleration_stats.h:15-16): inner nodes visited and
// triangle tests.
struct TraceCount {
    int64_t nodes = 0;
    int64_t tris = 0;
};
void TraceFinalOne(const FinalTree& t, const RayRec& r, uint32_t flags, HitRec& h, TraceCount* count = nullptr);
// TraceFinal with the traversal work summed over all rays.
void TraceFinalCounted(const FinalTree& t, const RayRec* rays, int64_t n, uint32_t flags, HitRec* hits, int threads,
                       TraceCount& total);
void TraceBrute(const float* tri_xyz, int64_t ntri, const RayRec* rays, int64_t n, uint32_t flags, HitRec* hits,
                int threads);
// The reference's brute force (src/acceleration/dummy_acceleration_structure.cc:21-79):
// Moller-Trumbore with the absolute |det| < 1e-6 rejection, closest hit t in (1e-6, t_max)
// with the first triangle winning a tie (strict <), any hit at t <= t_max (not <).
void TraceDummy(const float* tri_xyz, int64_t ntri, const RayRec* rays, int64_t n, uint32_t flags, HitRec* hits,
                int threads);

// ---------------------------------------------------------------------------------------
// Wide trees (wide_restate.cc): node i of width W is node i of the two-child tree opened
// greedily to W children; bytes: origin xyz (f32), meta (exponent byte per axis | count << 24),
// W child references, 6 planes of W bytes (lo x, hi x, lo y, hi y, lo z, hi z); 64 B for
// W = 4 (8 B zero tail), 96 B for W = 8. Child k's bound = fmaf(q, 2^(e - 127), origin).
// ---------------------------------------------------------------------------------------
struct WideTree {
    int width = 4;
    int64_t node_count = 0;
    std::vector<uint8_t> nodes;      // node_count * WideNodeBytes(width)
    std::vector<float> tri_ordered;  // as FinalTree::tri_ordered
};
int WideNodeBytes(int width);
void RefitAabb(const FinalTree& t, std::vector<FinalTree::Node>& out);
void Widen(const std::vector<FinalTree::Node>& nodes, int width, std::vector<uint8_t>& out);
void DecodeWideChild(const uint8_t* node, int width, int k, float lo[3], float hi[3]);
void TraceWideOne(const WideTree& t, const RayRec& r, uint32_t flags, HitRec& h, TraceCount* count = nullptr);
void TraceWideCounted(const WideTree& t, const RayRec* rays, int64_t n, uint32_t flags, HitRec* hits, int threads,
                      TraceCount& total);
double IntersectFp64(const float* tri_xyz9, const RayRec& r, double& b1, double& b2);

// ---------------------------------------------------------------------------------------
// Path tracing (diffuse, cosine-sampled, PCG32), forward and adjoint
// ---------------------------------------------------------------------------------------
struct Pcg32 {
    uint64_t state, inc;
};
void PcgSeed(Pcg32& r, uint64_t seed, uint64_t stream);
uint32_t PcgNext(Pcg32& r);
float PcgFloat(Pcg32& r);
void SinCos2Pi(float u, float& s, float& c);  // polynomial sin/cos of 2*pi*u, u in [0,1)

}  // namespace orc

namespace orc {
void CameraBasis(const float* cam, int w, int h, float* out12);

// Whitted ray tracing (cpu_raytracer.cc). Materials as src/scene/material.h:7-18.
constexpr int kWhittedMaxDepth = 8;
struct WMaterial {
    float diffuse[3], specular[3], transmittance[3], emission[3];
    float shininess, ior, r_ior, pad;
};
struct WParams {
    int W, H, max_depth, samples, cull;
    float background[3];
    float cam[11];
};
float PowPos(float x, float y);
float AreaLightSample01(uint32_t seed, uint32_t px, uint32_t py, uint32_t light_id, uint32_t sample_idx,
                        uint32_t dimension);
// The counters the reference's CpuRaytracer keeps (src/renderer/cpu_raytracer.cc:103,131,147,
// 223; acceleration stats from every Intersect / IntersectAny, sbvh.cc:61,148).
struct WStats {
    int64_t primary = 0;    // one per pixel
    int64_t secondary = 0;  // reflection and refraction rays
    int64_t shadow = 0;     // IsShadowed calls
    int64_t queries = 0;    // closest-hit and any-hit traces
    int64_t nodes = 0;      // inner nodes visited
    int64_t tris = 0;       // triangle tests
    void Add(const WStats& o) {
        primary += o.primary, secondary += o.secondary, shadow += o.shadow;
        queries += o.queries, nodes += o.nodes, tris += o.tris;
    }
};
void WhittedRender(const FinalTree& tree, const float* tri, const float* vn, int64_t n, const int32_t* mat_id,
                   const WMaterial* mats, const WParams& p, bool recursive, float* image, int threads,
                   WStats* stats = nullptr);
int TracePathVertices(const FinalTree& T, const float* tri_xyz, const int32_t* mat_id, const float* albedo,
                      const float* basis, float far, int x, int y, int W, int spp, int s, int depth, uint64_t seed,
                      int32_t* mats, Vec3* betas);
}  // namespace orc
