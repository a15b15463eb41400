// Host side of the C++ interface: the reference-shaped types and the GpuBvh / GpuPathTracer
// adapters over the C ABI. See dpl_host.h.
#include "dpl_host.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>

namespace diplodocus_amd {

float RayEpsilon(const Vec3& p, float t_hit) {
    const float p_scale = std::max({std::fabs(p.x), std::fabs(p.y), std::fabs(p.z), 1.0f});
    const float t_scale = std::max(t_hit, 1.0f);
    const float bias = std::max(1e-6f * p_scale, 1e-7f * t_scale);
    return std::clamp(bias, 1e-7f, 1e-3f);
}

AccelerationStructureType ParseAccelerationStructureType(std::string_view sv) {
    if (sv == "dummy") return AccelerationStructureType::kDummy;
    if (sv == "sbvh") return AccelerationStructureType::kSbvh;
    if (sv == "ploc") return AccelerationStructureType::kPloc;
    if (sv == "ploc_emc") return AccelerationStructureType::kPlocEmc;
    if (sv == "ploc_sobb") return AccelerationStructureType::kPlocSobb;
    if (sv == "ploc_emc_sobb") return AccelerationStructureType::kPlocEmcSobb;
    throw std::runtime_error("Invalid AccelerationStructureType: " + std::string(sv));
}

// ---- Context ----
Context::Context(int device) {
    const dpl_status s = dpl_context_create(device, &ctx_);
    if (s != DPL_OK) throw DplError(s, "dpl_context_create failed (no usable GPU?)");
}

Context::~Context() { dpl_context_destroy(ctx_); }

void Context::Check(dpl_status s) const {
    if (s != DPL_OK) throw DplError(s, std::string("dpl: ") + dpl_last_error(ctx_));
}

// ---- GpuBvh ----
GpuBvh::GpuBvh(std::shared_ptr<Context> ctx) : ctx_(ctx ? std::move(ctx) : std::make_shared<Context>(0)) {}

GpuBvh::~GpuBvh() { dpl_bvh_destroy(bvh_); }

void GpuBvh::Build(const AccelerationStructureConfig& config, Stats& stats, std::span<const Triangle> triangles) {
    dpl_build_params p{};
    const dpl_status s =
        dpl_build_params_from_type(static_cast<int32_t>(config.acceleration_structure_type), config.nn_search_radius,
                                   config.max_triangles_per_leaf, config.kdop_size, &p);
    if (s == DPL_ERR_UNSUPPORTED) throw std::invalid_argument("GpuBvh builds the PLOC acceleration structure types only");
    ctx_->Check(s);
    // the build reads positions only: n x 9 floats
    std::vector<float> xyz(triangles.size() * 9);
    for (size_t i = 0; i < triangles.size(); i++) {
        const Triangle& t = triangles[i];
        const Vec3* v[3] = {&t.v0.pos, &t.v1.pos, &t.v2.pos};
        for (int k = 0; k < 3; k++) {
            xyz[9 * i + 3 * k] = v[k]->x;
            xyz[9 * i + 3 * k + 1] = v[k]->y;
            xyz[9 * i + 3 * k + 2] = v[k]->z;
        }
    }
    dpl_build_stats bs{};
    if (bvh_) {
        ctx_->Check(dpl_bvh_rebuild(bvh_, xyz.data(), int64_t(triangles.size()), DPL_MEM_HOST, &p, &bs));
    } else {
        ctx_->Check(dpl_bvh_build(ctx_->get(), xyz.data(), int64_t(triangles.size()), DPL_MEM_HOST, &p, &bvh_, &bs));
    }
    triangles_ = triangles;
    build_stats_ = bs;
    wide_width_ = 0;  // a rebuild drops the wide form
    AccelerationStats& a = stats.accel_stats;
    a.build_time = bs.build_ms;
    a.node_count = bs.node_count;
    a.inner_node_count = bs.inner_node_count;
    a.leaf_node_count = bs.leaf_node_count;
    a.memory_consumption = int(std::min<int64_t>(bs.memory_consumption, INT32_MAX));
}

void GpuBvh::Widen(int width) {
    if (!bvh_) throw std::logic_error("GpuBvh: Build before Widen");
    if (width != 4 && width != 8) throw std::invalid_argument("GpuBvh::Widen: width must be 4 or 8");
    ctx_->Check(dpl_bvh_widen(bvh_, width, nullptr));
    wide_width_ = width;
}

std::vector<dpl_hit> GpuBvh::Trace(Stats& stats, std::span<const Ray> rays, uint32_t flags) const {
    if (!bvh_) throw std::logic_error("GpuBvh: Build before tracing");
    if (wide_width_) flags |= DPL_TRACE_WIDE;
    std::vector<dpl_ray> r(rays.size());
    for (size_t i = 0; i < rays.size(); i++) {
        const Ray& q = rays[i];
        r[i] = dpl_ray{q.origin.x, q.origin.y, q.origin.z, kEpsilon, q.dir.x, q.dir.y, q.dir.z, q.t_max};
    }
    std::vector<dpl_hit> h(rays.size());
    int64_t visits = 0, tests = 0;
    ctx_->Check(dpl_trace_counters(ctx_->get(), bvh_, r.data(), int64_t(r.size()), flags, h.data(), DPL_MEM_HOST,
                                   &visits, &tests));
    AccelerationStats& a = stats.accel_stats;
    a.query_count += int(rays.size());
    a.traversal_count += int(std::min<int64_t>(visits, INT32_MAX));
    a.intersection_count += int(std::min<int64_t>(tests, INT32_MAX));
    return h;
}

// RayHit as the reference fills it (src/acceleration/sbvh.cc:128-142).
void GpuBvh::FillHit(const Ray& ray, const dpl_hit& h, RayHit& out) const {
    const Triangle& tri = triangles_[size_t(h.prim)];
    out.t = h.t;
    out.triangle_id = h.prim;
    out.material_id = tri.material_id;
    out.pos = ray.At(h.t);
    out.b0 = 1.0f - h.b1 - h.b2;
    out.b1 = h.b1;
    out.b2 = h.b2;
    out.epsilon = RayEpsilon(out.pos, out.t);
    out.geom_normal = tri.geom_normal;
    out.normal = tri.geom_normal;
    if (tri.has_vertex_normals)
        out.normal = Normalize(tri.v0.normal * out.b0 + tri.v1.normal * out.b1 + tri.v2.normal * out.b2);
}

bool GpuBvh::Intersect(Stats& stats, const Ray& ray, RayHit& ray_hit, bool backface_culling) const {
    const auto h = Trace(stats, std::span<const Ray>(&ray, 1), backface_culling ? DPL_TRACE_CULL_BACK : 0u);
    if (h[0].prim < 0) return false;
    FillHit(ray, h[0], ray_hit);
    return true;
}

bool GpuBvh::IntersectAny(Stats& stats, const Ray& ray, bool backface_culling) const {
    const auto h = Trace(stats, std::span<const Ray>(&ray, 1),
                         DPL_TRACE_ANY | (backface_culling ? DPL_TRACE_CULL_BACK : 0u));
    return h[0].prim >= 0;
}

int64_t GpuBvh::IntersectBatch(Stats& stats, std::span<const Ray> rays, std::span<RayHit> hits,
                               bool backface_culling) const {
    if (hits.size() < rays.size()) throw std::invalid_argument("IntersectBatch: hits is shorter than rays");
    const auto h = Trace(stats, rays, backface_culling ? DPL_TRACE_CULL_BACK : 0u);
    int64_t count = 0;
    for (size_t i = 0; i < rays.size(); i++) {
        if (h[i].prim < 0) {
            hits[i] = RayHit{};
            continue;
        }
        FillHit(rays[i], h[i], hits[i]);
        count++;
    }
    return count;
}

void GpuBvh::IntersectAnyBatch(Stats& stats, std::span<const Ray> rays, std::span<uint8_t> occluded,
                               bool backface_culling) const {
    if (occluded.size() < rays.size()) throw std::invalid_argument("IntersectAnyBatch: occluded is too short");
    const auto h = Trace(stats, rays, DPL_TRACE_ANY | (backface_culling ? DPL_TRACE_CULL_BACK : 0u));
    for (size_t i = 0; i < rays.size(); i++) occluded[i] = h[i].prim >= 0 ? 1 : 0;
}

std::unique_ptr<AccelerationStructure> CreateGpuAccelerationStructure(const AccelerationStructureConfig& config,
                                                                      std::shared_ptr<Context> ctx) {
    switch (config.acceleration_structure_type) {
        case AccelerationStructureType::kPloc:
        case AccelerationStructureType::kPlocEmc:
        case AccelerationStructureType::kPlocSobb:
        case AccelerationStructureType::kPlocEmcSobb:
            return std::make_unique<GpuBvh>(std::move(ctx));
        default:
            throw std::invalid_argument("Unkown AccelerationStructureType for the GPU engine");
    }
}

// ---- GpuPathTracer ----
namespace {

void HipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw DplError(DPL_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// A device buffer owned for the duration of one call.
struct DeviceFloats {
    float* p = nullptr;
    explicit DeviceFloats(size_t n) { HipCheck(hipMalloc(&p, std::max<size_t>(n, 1) * 4), "hipMalloc"); }
    ~DeviceFloats() { (void)hipFree(p); }
    void Upload(std::span<const float> h) { HipCheck(hipMemcpy(p, h.data(), h.size() * 4, hipMemcpyHostToDevice), "upload"); }
    void Download(std::span<float> h) const {
        HipCheck(hipMemcpy(h.data(), p, h.size() * 4, hipMemcpyDeviceToHost), "download");
    }
};

}  // namespace

namespace {

// Device copies of host arrays for one call.
template <typename T>
struct DeviceArray {
    T* p = nullptr;
    explicit DeviceArray(const std::vector<T>& h) {
        HipCheck(hipMalloc(&p, std::max<size_t>(h.size(), 1) * sizeof(T)), "hipMalloc");
        if (!h.empty()) HipCheck(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
    }
    ~DeviceArray() { (void)hipFree(p); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
};

}  // namespace

std::vector<float> GpuWhittedRenderer::Render(const RenderConfig& cfg, const Camera& cam,
                                              std::span<const Triangle> triangles, std::span<const Material> materials,
                                              Stats* stats) const {
    const size_t n = triangles.size();
    std::vector<float> xyz(n * 9), vn(n * 9, 0.0f);
    std::vector<int32_t> mat(n), lights;
    bool any_vn = false;
    for (size_t i = 0; i < n; i++) {
        const Triangle& t = triangles[i];
        const Vertex* v[3] = {&t.v0, &t.v1, &t.v2};
        for (int k = 0; k < 3; k++) {
            xyz[9 * i + 3 * k] = v[k]->pos.x, xyz[9 * i + 3 * k + 1] = v[k]->pos.y, xyz[9 * i + 3 * k + 2] = v[k]->pos.z;
            if (t.has_vertex_normals) {
                vn[9 * i + 3 * k] = v[k]->normal.x, vn[9 * i + 3 * k + 1] = v[k]->normal.y;
                vn[9 * i + 3 * k + 2] = v[k]->normal.z;
            }
        }
        any_vn |= t.has_vertex_normals;
        if (t.material_id < 0 || size_t(t.material_id) >= materials.size())
            throw std::invalid_argument("material_id out of range");
        mat[i] = t.material_id;
        if (!(materials[size_t(t.material_id)].emission == Vec3{})) lights.push_back(int32_t(i));
    }
    std::vector<dpl_material> dm(materials.size());
    for (size_t m = 0; m < materials.size(); m++) {
        const Material& a = materials[m];
        dm[m] = dpl_material{{a.diffuse.x, a.diffuse.y, a.diffuse.z},
                             {a.specular.x, a.specular.y, a.specular.z},
                             {a.transmittance.x, a.transmittance.y, a.transmittance.z},
                             {a.emission.x, a.emission.y, a.emission.z},
                             a.shininess,
                             a.ior,
                             a.r_ior,
                             0.0f};
    }
    dpl_whitted_params wp{};
    wp.width = cfg.width, wp.height = cfg.height, wp.max_depth = cfg.max_depth;
    wp.area_light_sample_cnt = cfg.area_light_sample_cnt;
    wp.backface_culling = cfg.backface_culling ? 1 : 0;
    wp.background[0] = cfg.background_color.x, wp.background[1] = cfg.background_color.y;
    wp.background[2] = cfg.background_color.z;
    const float c[11] = {cam.pos.x, cam.pos.y, cam.pos.z, cam.dir.x, cam.dir.y, cam.dir.z,
                         cam.up.x,  cam.up.y,  cam.up.z,  cam.fov,   cam.far};
    std::memcpy(wp.camera, c, sizeof(c));
    DeviceArray<float> d_xyz(xyz), d_vn(any_vn ? vn : std::vector<float>{});
    DeviceArray<int32_t> d_mat(mat), d_lights(lights);
    DeviceArray<dpl_material> d_mats(dm);
    DeviceFloats img(size_t(cfg.width) * cfg.height * 3);
    const auto& ctx = bvh_.context();
    if (stats) {
        dpl_whitted_stats ws{};
        ctx->Check(dpl_render_whitted_stats(ctx->get(), bvh_.handle(), &wp, d_xyz.p, any_vn ? d_vn.p : nullptr,
                                            d_mat.p, d_mats.p, int32_t(dm.size()),
                                            lights.empty() ? nullptr : d_lights.p, int32_t(lights.size()), img.p, &ws));
        auto add = [](int& dst, int64_t v) { dst = int(std::min<int64_t>(int64_t(dst) + v, INT32_MAX)); };
        RaytracingStats& r = stats->rt_stats;
        r.raytracing_time = ws.raytracing_ms * 1e-3;
        add(r.primary_ray_count, ws.primary_ray_count);
        add(r.secondary_ray_count, ws.secondary_ray_count);
        add(r.shadow_ray_count, ws.shadow_ray_count);
        AccelerationStats& a = stats->accel_stats;
        add(a.query_count, ws.query_count);
        add(a.traversal_count, ws.traversal_count);
        add(a.intersection_count, ws.intersection_count);
    } else {
        ctx->Check(dpl_render_whitted(ctx->get(), bvh_.handle(), &wp, d_xyz.p, any_vn ? d_vn.p : nullptr, d_mat.p,
                                      d_mats.p, int32_t(dm.size()), lights.empty() ? nullptr : d_lights.p,
                                      int32_t(lights.size()), img.p));
    }
    ctx->Check(dpl_context_synchronize(ctx->get()));
    std::vector<float> out(size_t(cfg.width) * cfg.height * 3);
    img.Download(out);
    return out;
}

GpuPathTracer::GpuPathTracer(const GpuBvh& bvh, std::span<const Triangle> triangles, int material_count)
    : bvh_(bvh), material_count_(material_count) {
    std::vector<int32_t> ids(triangles.size());
    for (size_t i = 0; i < triangles.size(); i++) {
        ids[i] = triangles[i].material_id;
        if (ids[i] < 0 || ids[i] >= material_count) throw std::invalid_argument("material_id out of range");
    }
    HipCheck(hipMalloc(&mat_id_dev_, std::max<size_t>(ids.size(), 1) * 4), "hipMalloc");
    HipCheck(hipMemcpy(mat_id_dev_, ids.data(), ids.size() * 4, hipMemcpyHostToDevice), "upload");
}

GpuPathTracer::~GpuPathTracer() { (void)hipFree(mat_id_dev_); }

dpl_render_params GpuPathTracer::Params(const PathTraceConfig& cfg, const Camera& cam) const {
    dpl_render_params rp{};
    rp.width = cfg.width;
    rp.height = cfg.height;
    rp.spp = cfg.spp;
    rp.max_depth = cfg.max_depth;
    rp.seed = cfg.seed;
    const float c[11] = {cam.pos.x, cam.pos.y, cam.pos.z, cam.dir.x, cam.dir.y, cam.dir.z,
                         cam.up.x,  cam.up.y,  cam.up.z,  cam.fov,   cam.far};
    std::memcpy(rp.camera, c, sizeof(c));
    rp.material_count = material_count_;
    return rp;
}

std::vector<float> GpuPathTracer::Render(const PathTraceConfig& cfg, const Camera& cam, std::span<const float> albedo,
                                         std::span<const float> emission) const {
    const size_t m3 = size_t(material_count_) * 3, px = size_t(cfg.width) * cfg.height * 3;
    if (albedo.size() != m3 || emission.size() != m3) throw std::invalid_argument("material arrays must be M x 3");
    DeviceFloats a(m3), e(m3), img(px);
    a.Upload(albedo);
    e.Upload(emission);
    const dpl_render_params rp = Params(cfg, cam);
    bvh_.context()->Check(dpl_render_forward(bvh_.context()->get(), bvh_.handle(), &rp, mat_id_dev_, a.p, e.p, img.p));
    bvh_.context()->Check(dpl_context_synchronize(bvh_.context()->get()));
    std::vector<float> out(px);
    img.Download(out);
    return out;
}

void GpuPathTracer::Adjoint(const PathTraceConfig& cfg, const Camera& cam, std::span<const float> albedo,
                            std::span<const float> emission, std::span<const float> d_image,
                            std::span<float> grad_albedo, std::span<float> grad_emission) const {
    const size_t m3 = size_t(material_count_) * 3, px = size_t(cfg.width) * cfg.height * 3;
    if (albedo.size() != m3 || emission.size() != m3 || grad_albedo.size() != m3 || grad_emission.size() != m3)
        throw std::invalid_argument("material arrays must be M x 3");
    if (d_image.size() != px) throw std::invalid_argument("d_image must be height x width x 3");
    DeviceFloats a(m3), e(m3), di(px), ga(m3), ge(m3);
    a.Upload(albedo);
    e.Upload(emission);
    di.Upload(d_image);
    const dpl_render_params rp = Params(cfg, cam);
    bvh_.context()->Check(
        dpl_render_adjoint(bvh_.context()->get(), bvh_.handle(), &rp, mat_id_dev_, a.p, e.p, di.p, ga.p, ge.p));
    bvh_.context()->Check(dpl_context_synchronize(bvh_.context()->get()));
    ga.Download(grad_albedo);
    ge.Download(grad_emission);
}

}  // namespace diplodocus_amd
