// C++20 host interface of the engine over the C ABI (include/dpl/dpl.h). It keeps the
// reference's plug-in shape so that a renderer written against it switches by one factory
// case:
//   * scene and ray types: Vec3 (src/util/vec3.h), Vertex (src/scene/vertex.h:7-12),
//     Triangle (src/scene/triangle.h:12-20), Ray + RayEpsilon (src/scene/ray.h:7-24),
//     RayHit (src/scene/ray_hit.h:8-17), Camera (src/scene/camera.h:7-13);
//   * configuration: AccelerationStructureType / AccelerationStructureConfig
//     (src/config/acceleration_structure_config.h:9-29);
//   * statistics: AccelerationStats, RaytracingStats, Stats (src/stats/*.h);
//   * the plug-in: AccelerationStructure::{Build, Intersect, IntersectAny}
//     (src/acceleration/acceleration_structure.h:12-19) implemented by GpuBvh, plus batched
//     IntersectBatch / IntersectAnyBatch that trace a whole ray span in one launch.
// Errors: dpl failures throw DplError (a std::runtime_error); an acceleration type this
// library does not build (kDummy, kSbvh) throws std::invalid_argument, as the reference's
// factory does for unknown types (src/acceleration/acceleration_structure_factory.cc:16-17).
#pragma once

#include <cmath>
#include <cstdint>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "dpl/dpl.h"

namespace diplodocus_amd {

constexpr float kEpsilon = 1e-6f;  // the reference's implicit ray t_min (src/util/util.h)

struct Vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    Vec3 operator+(const Vec3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vec3 operator-(const Vec3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vec3 operator*(float s) const { return {x * s, y * s, z * s}; }
    Vec3 operator-() const { return {-x, -y, -z}; }
    bool operator==(const Vec3& o) const = default;
};
inline float Dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 Cross(const Vec3& a, const Vec3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float Length(const Vec3& v) { return std::sqrt(Dot(v, v)); }
inline Vec3 Normalize(const Vec3& v) {
    const float l = Length(v);
    return l <= 1e-6f ? Vec3{} : Vec3{v.x / l, v.y / l, v.z / l};
}

struct Vertex {
    Vec3 pos;
    Vec3 normal;
};

struct Triangle {
    Vertex v0, v1, v2;
    Vec3 geom_normal;
    int32_t material_id = 0;
    bool has_vertex_normals = false;
};

struct Ray {
    Vec3 origin;
    Vec3 dir;
    float t_max = 0.0f;
    Vec3 At(float t) const { return origin + dir * t; }
};

// Offset scale of a hit point (src/scene/ray.h:15-24).
float RayEpsilon(const Vec3& hit_pos, float t_hit);

struct RayHit {
    Vec3 pos;
    Vec3 normal;
    Vec3 geom_normal;
    float b0 = 0.0f, b1 = 0.0f, b2 = 0.0f;
    float t = 0.0f;
    int triangle_id = -1;
    int material_id = -1;
    float epsilon = kEpsilon;
};

// src/scene/material.h:7-18
struct Material {
    std::string name;
    Vec3 ambient, diffuse, specular, transmittance, emission;
    float shininess = 0.0f;
    float ior = 1.0f;
    float r_ior = 1.0f;  // 1 / ior, as the reference's OBJ loader sets it
    float dissolve = 1.0f;
};

enum class RendererType { kCpu, kGpu };

// src/config/render_config.h:17-27
struct RenderConfig {
    RendererType renderer_type = RendererType::kCpu;
    int width = 800;
    int height = 600;
    Vec3 background_color;
    bool backface_culling = false;
    int max_depth = 4;
    int area_light_sample_cnt = 8;
    int pixel_sample_cnt = 4;
    int seed = 42;
};

struct Camera {
    Vec3 pos{0.0f, 0.0f, 0.0f};
    Vec3 dir{0.0f, 0.0f, 1.0f};
    Vec3 up{0.0f, 1.0f, 0.0f};
    float fov = 0.75f;  // radians
    float far = 100.0f;
};

enum class AccelerationStructureType { kDummy, kSbvh, kPloc, kPlocEmc, kPlocSobb, kPlocEmcSobb };

// "dummy", "sbvh", "ploc", "ploc_emc", "ploc_sobb", "ploc_emc_sobb"; throws std::runtime_error
// on anything else.
AccelerationStructureType ParseAccelerationStructureType(std::string_view sv);

struct AccelerationStructureConfig {
    AccelerationStructureType acceleration_structure_type = AccelerationStructureType::kDummy;
    int max_depth = 32;
    int nn_search_radius = 10;
    int kdop_size = 32;
    int max_triangles_per_leaf = 4;
    int bin_count = 16;
};

struct AccelerationStats {
    double build_time = 0.0;  // ms
    int node_count = 0;
    int inner_node_count = 0;
    int leaf_node_count = 0;
    int memory_consumption = 0;
    int query_count = 0;
    int intersection_count = 0;
    int traversal_count = 0;
};

struct RaytracingStats {
    double raytracing_time = 0.0;
    int primary_ray_count = 0;
    int secondary_ray_count = 0;
    int shadow_ray_count = 0;
};

struct Stats {
    RaytracingStats rt_stats;
    AccelerationStats accel_stats;
};

class AccelerationStructure {
   public:
    virtual ~AccelerationStructure() = default;
    virtual void Build(const AccelerationStructureConfig& acceleration_structure_config, Stats& stats,
                       std::span<const Triangle> triangles) = 0;
    virtual bool Intersect(Stats& stats, const Ray& ray, RayHit& ray_hit, bool backface_culling) const = 0;
    virtual bool IntersectAny(Stats& stats, const Ray& ray, bool backface_culling) const = 0;
};

class DplError : public std::runtime_error {
   public:
    DplError(dpl_status s, const std::string& what) : std::runtime_error(what), status(s) {}
    dpl_status status;
};

// One device and one stream (dpl_context).
class Context {
   public:
    explicit Context(int device = 0);
    ~Context();
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    dpl_context* get() const { return ctx_; }
    void Check(dpl_status s) const;  // throws DplError with dpl_last_error on failure

   private:
    dpl_context* ctx_ = nullptr;
};

// The GPU PLOC BVH behind the reference's plug-in interface. Like the reference's SBvh it
// keeps the triangle span given to Build (the scene must outlive the BVH) to fill RayHit's
// material, normals and barycentrics. Intersect / IntersectAny trace one ray per call (a
// correctness path: each call is one launch and one synchronisation); IntersectBatch /
// IntersectAnyBatch trace a span of rays in one launch and count traversal work.
class GpuBvh final : public AccelerationStructure {
   public:
    explicit GpuBvh(std::shared_ptr<Context> ctx = nullptr);
    ~GpuBvh() override;
    GpuBvh(const GpuBvh&) = delete;
    GpuBvh& operator=(const GpuBvh&) = delete;

    void Build(const AccelerationStructureConfig& config, Stats& stats, std::span<const Triangle> triangles) override;
    bool Intersect(Stats& stats, const Ray& ray, RayHit& ray_hit, bool backface_culling) const override;
    bool IntersectAny(Stats& stats, const Ray& ray, bool backface_culling) const override;

    // Returns the number of rays that hit; misses get triangle_id = -1.
    int64_t IntersectBatch(Stats& stats, std::span<const Ray> rays, std::span<RayHit> hits,
                           bool backface_culling) const;
    // occluded[i] = 1 when ray i hits anything in (kEpsilon, t_max).
    void IntersectAnyBatch(Stats& stats, std::span<const Ray> rays, std::span<uint8_t> occluded,
                           bool backface_culling) const;

    // Adds the wide form (4 or 8 children, quantized child boxes; dpl_bvh_widen): every trace
    // until the next Build traverses it, with the same closest hits. No counterpart in the
    // reference.
    void Widen(int width = 4);
    int wide_width() const { return wide_width_; }

    const dpl_build_stats& build_stats() const { return build_stats_; }
    dpl_bvh* handle() const { return bvh_; }
    const std::shared_ptr<Context>& context() const { return ctx_; }

   private:
    void FillHit(const Ray& ray, const dpl_hit& h, RayHit& out) const;
    std::vector<dpl_hit> Trace(Stats& stats, std::span<const Ray> rays, uint32_t flags) const;

    std::shared_ptr<Context> ctx_;
    dpl_bvh* bvh_ = nullptr;
    std::span<const Triangle> triangles_;
    dpl_build_stats build_stats_{};
    int wide_width_ = 0;
};

// The factory case a renderer adds for the PLOC types (kPloc, kPlocEmc, kPlocSobb,
// kPlocEmcSobb); throws std::invalid_argument for the others.
std::unique_ptr<AccelerationStructure> CreateGpuAccelerationStructure(const AccelerationStructureConfig& config,
                                                                      std::shared_ptr<Context> ctx = nullptr);

// The reference's Whitted CpuRaytracer (src/renderer/cpu_raytracer.cc:50-288) on the GPU over
// a GpuBvh built from the same triangles (dpl_render_whitted): area lights are the emissive
// triangles in triangle order, as the reference's loader makes them. Returns the framebuffer
// (row-major, height x width x 3, src/framebuffer/framebuffer.cc:47-52). With stats, adds the
// render's counters as the reference's renderer does (cpu_raytracer.cc:84-88,103,131,147,223):
// primary / secondary / shadow rays, raytracing_time (s, device time), and the
// acceleration-structure queries with their traversal work (dpl_render_whitted_stats).
class GpuWhittedRenderer {
   public:
    explicit GpuWhittedRenderer(const GpuBvh& bvh) : bvh_(bvh) {}
    std::vector<float> Render(const RenderConfig& cfg, const Camera& cam, std::span<const Triangle> triangles,
                              std::span<const Material> materials, Stats* stats = nullptr) const;

   private:
    const GpuBvh& bvh_;
};

// Diffuse path tracing over a GpuBvh with per-material albedo and emission
// (dpl_render_forward), and the adjoint path replay that gives the image-loss gradient with
// respect to them (dpl_render_adjoint). Materials are indexed by Triangle::material_id.
struct PathTraceConfig {
    int width = 512, height = 512, spp = 4, max_depth = 4;
    uint64_t seed = 1;
};

class GpuPathTracer {
   public:
    GpuPathTracer(const GpuBvh& bvh, std::span<const Triangle> triangles, int material_count);
    ~GpuPathTracer();
    GpuPathTracer(const GpuPathTracer&) = delete;
    GpuPathTracer& operator=(const GpuPathTracer&) = delete;

    // albedo / emission: material_count x 3 floats. Returns height x width x 3 floats.
    std::vector<float> Render(const PathTraceConfig& cfg, const Camera& cam, std::span<const float> albedo,
                              std::span<const float> emission) const;
    // d_image: dL/dimage (height x width x 3). Writes material_count x 3 gradients each.
    void Adjoint(const PathTraceConfig& cfg, const Camera& cam, std::span<const float> albedo,
                 std::span<const float> emission, std::span<const float> d_image, std::span<float> grad_albedo,
                 std::span<float> grad_emission) const;

   private:
    dpl_render_params Params(const PathTraceConfig& cfg, const Camera& cam) const;
    const GpuBvh& bvh_;
    int material_count_;
    int32_t* mat_id_dev_ = nullptr;  // per original triangle, device
};

}  // namespace diplodocus_amd
