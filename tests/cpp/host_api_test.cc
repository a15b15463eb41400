// Drives the C++ host interface the way the reference's renderer drives an
// AccelerationStructure (src/renderer/cpu_raytracer.cc:56-57,112,236): build through the
// factory, per-ray Intersect / IntersectAny, plus the batched calls, and the path tracer.
//   host_api_test --cpu [--scenes DIR] [--configs DIR]   interface checks that need no GPU
//   host_api_test [--scenes DIR]         the GPU checks; hits are compared with the oracle's
//                                        brute force, images with the oracle's renderer
// The scenes DIR holds the OBJ scene fixtures (CornellBox-Original / CornellBox-Sphere), the
// configs DIR the JSON configuration fixtures; their checks are skipped without them.
// Prints one line per check group and "host_api_test: OK" at the end; exits non-zero on the
// first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include <filesystem>
#include <sstream>

#include "config.h"
#include "dpl_host.h"
#include "renderer.h"
#include "scene_loader.h"

extern "C" void orc_trace_brute(const float* tri, int64_t ntri, const void* rays, int64_t n, uint32_t flags, void* hits,
                                int threads);
extern "C" void* orc_ploc_build(const float* tri, int64_t n, int morton_type, int radius, int max_leaf, int bv_type,
                                int kdop_size, double* build_ms);
extern "C" void orc_ploc_free(void* h);
extern "C" void orc_whitted(void* ploc, const float* tri_xyz, const float* vertex_normals, int64_t n,
                            const int32_t* mat_id, const float* mats, int W, int H, int max_depth, int samples, int cull,
                            const float* background, const float* cam, int recursive, float* image, int threads,
                            int64_t* counts);

using namespace diplodocus_amd;

namespace {

int g_failures = 0;
std::string g_scene_dir;
std::string g_config_dir;


void Expect(bool ok, const std::string& what) {
    if (!ok) {
        std::printf("FAILED: %s\n", what.c_str());
        std::fflush(stdout);
        g_failures++;
    }
}

struct Lcg {
    uint64_t s;
    float Next() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return float((s >> 40) & 0xFFFFFF) / 16777216.0f;
    }
};

// A closed box (12 triangles, material 0..5 by face) around a seeded soup of small
// triangles (material 6), some with vertex normals.
std::vector<Triangle> MakeScene(int n_soup, uint64_t seed) {
    std::vector<Triangle> tris;
    auto quad = [&](Vec3 a, Vec3 b, Vec3 c, Vec3 d, int mat) {
        for (int k = 0; k < 2; k++) {
            Triangle t;
            t.v0.pos = a;
            t.v1.pos = k ? c : b;
            t.v2.pos = k ? d : c;
            t.geom_normal = Normalize(Cross(t.v1.pos - t.v0.pos, t.v2.pos - t.v0.pos));
            t.material_id = mat;
            tris.push_back(t);
        }
    };
    const float L = 0.0f, H = 1.0f;
    quad({L, L, L}, {H, L, L}, {H, L, H}, {L, L, H}, 0);  // floor
    quad({L, H, L}, {L, H, H}, {H, H, H}, {H, H, L}, 1);  // ceiling
    quad({L, L, H}, {H, L, H}, {H, H, H}, {L, H, H}, 2);  // back
    quad({L, L, L}, {L, L, H}, {L, H, H}, {L, H, L}, 3);  // left
    quad({H, L, L}, {H, H, L}, {H, H, H}, {H, L, H}, 4);  // right
    quad({L, L, L}, {L, H, L}, {H, H, L}, {H, L, L}, 5);  // front (the camera is inside)
    Lcg r{seed};
    for (int i = 0; i < n_soup; i++) {
        const Vec3 c{0.2f + 0.6f * r.Next(), 0.2f + 0.6f * r.Next(), 0.2f + 0.6f * r.Next()};
        Triangle t;
        Vertex* v[3] = {&t.v0, &t.v1, &t.v2};
        for (int k = 0; k < 3; k++)
            v[k]->pos = c + Vec3{0.04f * (r.Next() - 0.5f), 0.04f * (r.Next() - 0.5f), 0.04f * (r.Next() - 0.5f)};
        t.geom_normal = Normalize(Cross(t.v1.pos - t.v0.pos, t.v2.pos - t.v0.pos));
        t.material_id = 6;
        if (i % 3 == 0) {
            t.has_vertex_normals = true;
            for (int k = 0; k < 3; k++) v[k]->normal = Normalize(v[k]->pos - Vec3{0.5f, 0.5f, 0.5f});
        }
        tris.push_back(t);
    }
    return tris;
}

std::vector<Ray> CameraRays(int w, int h) {
    // src/renderer/cpu_raytracer.cc:62-73 with pos (0.5, 0.5, 0.05) looking along +z
    const Vec3 pos{0.5f, 0.5f, 0.05f}, dir{0, 0, 1}, up{0, 1, 0};
    const float fov = 1.2f;
    const Vec3 right = Normalize(Cross(dir, up));
    const float gw = float(2.0 * std::tan(fov / 2.0f)), gh = gw * (float(h) / float(w));
    const Vec3 qw = right * (gw / float(w - 1)), qh = -up * (gh / float(h - 1));
    const Vec3 p00 = dir - right * (gw / 2) + up * (gh / 2);
    std::vector<Ray> rays;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) rays.push_back(Ray{pos, Normalize(p00 + qw * float(x) + qh * float(y)), 100.0f});
    return rays;
}

// JSON configuration (the reference's keys and defaults) and the stats exporters.
void ConfigChecks() {
    JsonConfigLoader loader;
    if (!g_config_dir.empty()) {
        // the reference's own sample configuration (sample_config.json)
        const auto c = loader.Load(std::filesystem::path(g_config_dir) / "sample_config.json");
        Expect(c.has_value(), "config: sample_config.json loads: " + JsonConfigLoader::LastError());
        if (c) {
            const RenderConfig& r = c->render_config;
            Expect(r.renderer_type == RendererType::kCpu && r.width == 512 && r.height == 512 && r.max_depth == 8 &&
                       r.area_light_sample_cnt == 8 && r.background_color == Vec3{} && r.pixel_sample_cnt == 4 &&
                       r.seed == 42 && !r.backface_culling,
                   "config: render section and defaults");
            Expect(c->acceleration_structure_config.acceleration_structure_type == AccelerationStructureType::kSbvh &&
                       c->acceleration_structure_config.nn_search_radius == 10,
                   "config: acceleration structure section and defaults");
            Expect(c->scene_load_config.name == "CornellBox-Bunny" && c->scene_load_config.dirpath == "res",
                   "config: scene section");
            Expect(c->image_export_config.filepath == "output.ppm" &&
                       c->stats_export_config.stats_export_format == StatsExportFormat::kConsole,
                   "config: image and stats sections");
        }
        const auto f = loader.Load(std::filesystem::path(g_config_dir) / "full_config.json");
        Expect(f.has_value(), "config: full_config.json loads: " + JsonConfigLoader::LastError());
        if (f) {
            const RenderConfig& r = f->render_config;
            Expect(r.renderer_type == RendererType::kGpu && r.width == 1024 && r.height == 768 &&
                       r.background_color == Vec3{0.25f, 0.5f, 0.1f} && r.backface_culling && r.max_depth == 6 &&
                       r.area_light_sample_cnt == 16 && r.pixel_sample_cnt == 4 && r.seed == 7,
                   "config: numbers convert, null keeps the default, Vec3 reads three numbers");
            const AccelerationStructureConfig& a = f->acceleration_structure_config;
            Expect(a.acceleration_structure_type == AccelerationStructureType::kPlocEmcSobb && a.max_depth == 48 &&
                       a.max_triangles_per_leaf == 8 && a.nn_search_radius == 32 && a.kdop_size == 18 &&
                       a.bin_count == 32,
                   "config: acceleration keys (a repeated key keeps its last value)");
            Expect(f->scene_load_config.dirpath == "scenes/obj" && !f->scene_load_config.triangulate &&
                       f->stats_export_config.stats_export_format == StatsExportFormat::kCsv &&
                       f->stats_export_config.filepath == "stats.csv" &&
                       f->image_export_config.filepath == "renders/cornell.ppm",
                   "config: strings with escapes, enums, paths");
        }
    }
    Expect(!loader.Load("/nonexistent/config.json").has_value(), "config: a missing file fails");
    const char* bad[] = {"", "{", "{\"render\": {\"width\": 01}}", "{\"render\": {\"width\": \"800\"}}",
                         "{\"render\": {\"renderer_type\": \"vulkan\"}}", "{\"render\": {}} x",
                         "{\"render\": {\"backface_culling\": 1}}", "{\"render\": {\"background_color\": [1, 2]}}",
                         "{\"a\": [1, 2,]}", "// c\n{}", "{\"s\": \"\\ud800\"}"};
    for (const char* b : bad) Expect(!loader.LoadFromString(b).has_value(), std::string("config: rejects ") + b);
    const auto d = loader.LoadFromString("[1, 2, 3]");
    Expect(d.has_value() && d->render_config.width == 800, "config: a non-object document keeps every default");
    const auto e = loader.LoadFromString("{\"render\": 5, \"acceleration_structure\": {\"kdop_size\": 14.0}}");
    Expect(e.has_value() && e->render_config.width == 800 && e->acceleration_structure_config.kdop_size == 14,
           "config: a non-object section keeps its defaults");

    // stats exporters
    Stats st;
    st.rt_stats = RaytracingStats{1.25, 262144, 1000, 2000000};
    st.accel_stats = AccelerationStats{4.5, 2047, 1023, 1024, 65504, 100, 450, 1234};
    std::ostringstream con;
    CreateStatsExporter(StatsExportConfig{}, &con)->ExportOne(StatsExportConfig{}, st);
    const std::string want =
        "########################################\n#                                      #\n"
        "#              Statistics              #\n#                                      #\n"
        "########################################\n\n"
        "----------------------------------------\n               Ray tracing              \n"
        "----------------------------------------\n"
        " Ray tracing time:              1.25 s\n Primary ray count:             262144\n"
        " Secondary ray count:           1000\n Shadow ray count:              2000000\n\n"
        "----------------------------------------\n               Acceleration             \n"
        "----------------------------------------\n# Construction:                         \n"
        " Build time:                    4.5 ms\n Node count:                    2047\n"
        " Inner node count:              1023\n Leaf node count:               1024\n"
        " Memory consumption:            65.5 KB\n\n# Traversal:                             \n"
        " Query count:                   100\n Intersection count:            450\n"
        " Intersections per query (avg): 4.5\n Traversal count:               1234\n"
        " Traversals per query:          12.3\n\n########################################\n";
    Expect(con.str() == want, "stats: console report");
    std::ostringstream csv;
    StatsExportConfig sc;
    sc.stats_export_format = StatsExportFormat::kCsv;
    CreateStatsExporter(sc, &csv)->ExportOne(sc, st);
    Expect(csv.str() == CsvStatsExporter::Header() + "\n1.25,262144,1000,2000000,4.5,2047,1023,1024,65504,100,450,1234\n",
           "stats: CSV header and row");
    const std::filesystem::path tmp = std::filesystem::temp_directory_path() / "dpl_host_api_test_stats.csv";
    std::filesystem::remove(tmp);
    sc.filepath = tmp;
    CsvStatsExporter().ExportOne(sc, st);
    CsvStatsExporter().ExportOne(sc, st);
    std::ifstream in(tmp);
    std::string l0, l1, l2, l3;
    std::getline(in, l0), std::getline(in, l1), std::getline(in, l2);
    Expect(l0 == CsvStatsExporter::Header() && l1 == l2 && l1 == CsvStatsExporter::Row(st) && !std::getline(in, l3),
           "stats: CSV file gets one header and a row per export");
    std::filesystem::remove(tmp);
}

void CpuChecks() {
    Expect(ParseAccelerationStructureType("ploc_emc_sobb") == AccelerationStructureType::kPlocEmcSobb, "parse");
    Expect(ParseAccelerationStructureType("sbvh") == AccelerationStructureType::kSbvh, "parse sbvh");
    bool threw = false;
    try {
        ParseAccelerationStructureType("bvh8");
    } catch (const std::runtime_error&) {
        threw = true;
    }
    Expect(threw, "parse rejects unknown names");
    const AccelerationStructureConfig cfg;
    Expect(cfg.nn_search_radius == 10 && cfg.kdop_size == 32 && cfg.max_triangles_per_leaf == 4 &&
               cfg.bin_count == 16 && cfg.max_depth == 32,
           "config defaults");
    // RayEpsilon: clamp(max(1e-6 max(|p|inf, 1), 1e-7 max(t, 1)), 1e-7, 1e-3)
    Expect(RayEpsilon({0, 0, 0}, 0.5f) == 1e-6f, "epsilon at the origin");
    Expect(RayEpsilon({-300.0f, 2, 1}, 1.0f) == std::clamp(1e-6f * 300.0f, 1e-7f, 1e-3f), "epsilon scales with |p|");
    Expect(RayEpsilon({1e6f, 0, 0}, 1.0f) == 1e-3f, "epsilon upper clamp");
    Expect(RayEpsilon({0, 0, 0}, 5e3f) == 1e-7f * 5e3f, "epsilon scales with t");
    threw = false;
    try {
        AccelerationStructureConfig c;
        c.acceleration_structure_type = AccelerationStructureType::kSbvh;
        CreateGpuAccelerationStructure(c, nullptr);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    Expect(threw, "factory rejects the CPU structure types");
    if (!g_scene_dir.empty()) {
        // the reference's scene path: factory, Load(<dir>/<name>.meta + .obj)
        SceneLoadConfig sc;
        sc.dirpath = g_scene_dir;
        sc.name = "CornellBox-Original";
        const auto scene = CreateSceneLoader(sc)->Load(sc);
        Expect(scene.has_value(), "scene loader: " + ObjSceneLoader::LastError());
        if (scene) {
            Expect(scene->Triangles().size() == 34 && scene->Materials().size() == 8 &&
                       scene->AreaLights().size() == 2 && scene->PointLights().size() == 1,
                   "scene loader: Cornell box counts");
            const Camera& c = scene->GetCamera();
            Expect(c.pos == Vec3{278.0f, 273.0f, -1000.0f} && c.fov == 0.6f && c.far == 2000.0f,
                   "scene loader: camera from the .meta file");
            const AreaLight& al = scene->AreaLights()[0];
            Expect(scene->Materials()[size_t(scene->Triangles()[size_t(al.triangle_id)].material_id)].name == "light" &&
                       al.color == Vec3{17.0f, 12.0f, 4.0f} && std::fabs(al.surface_area - 6825.0f) < 1e-2f,
                   "scene loader: area light");
        }
        sc.name = "no-such-scene";
        Expect(!ObjSceneLoader().Load(sc).has_value() && !ObjSceneLoader::LastError().empty(),
               "scene loader: a missing scene fails with a reason");
        Expect(ParseSceneLoadFormat("obj") == SceneLoadFormat::kObj, "scene load format");
    }
    ConfigChecks();
    std::printf("cpu checks: %s\n", g_failures ? "failed" : "ok");
}

// The oracle's restatement of the reference's Whitted renderer over the same scene.
std::vector<float> OracleWhitted(const RenderConfig& rc, const Camera& cam, std::span<const Triangle> tris,
                                 std::span<const Material> mats) {
    std::vector<float> mrows(16 * mats.size(), 0.0f), vn(tris.size() * 9, 0.0f), xyz(tris.size() * 9);
    std::vector<int32_t> mid(tris.size());
    for (size_t m = 0; m < mats.size(); m++) {
        const Vec3 f[4] = {mats[m].diffuse, mats[m].specular, mats[m].transmittance, mats[m].emission};
        for (int k = 0; k < 4; k++) mrows[16 * m + 3 * k] = f[k].x, mrows[16 * m + 3 * k + 1] = f[k].y, mrows[16 * m + 3 * k + 2] = f[k].z;
        mrows[16 * m + 12] = mats[m].shininess, mrows[16 * m + 13] = mats[m].ior, mrows[16 * m + 14] = mats[m].r_ior;
    }
    for (size_t i = 0; i < tris.size(); i++) {
        mid[i] = tris[i].material_id;
        const Vertex* v[3] = {&tris[i].v0, &tris[i].v1, &tris[i].v2};
        for (int k = 0; k < 3; k++) {
            xyz[9 * i + 3 * k] = v[k]->pos.x, xyz[9 * i + 3 * k + 1] = v[k]->pos.y, xyz[9 * i + 3 * k + 2] = v[k]->pos.z;
            if (tris[i].has_vertex_normals)
                vn[9 * i + 3 * k] = v[k]->normal.x, vn[9 * i + 3 * k + 1] = v[k]->normal.y, vn[9 * i + 3 * k + 2] = v[k]->normal.z;
        }
    }
    void* ploc = orc_ploc_build(xyz.data(), int64_t(tris.size()), 0, 16, 4, 0, 32, nullptr);
    std::vector<float> ref(size_t(rc.width) * rc.height * 3);
    const float bg[3] = {rc.background_color.x, rc.background_color.y, rc.background_color.z};
    const float camv[11] = {cam.pos.x, cam.pos.y, cam.pos.z, cam.dir.x, cam.dir.y, cam.dir.z,
                            cam.up.x,  cam.up.y,  cam.up.z,  cam.fov,   cam.far};
    orc_whitted(ploc, xyz.data(), vn.data(), int64_t(tris.size()), mid.data(), mrows.data(), rc.width, rc.height,
                rc.max_depth, rc.area_light_sample_cnt, rc.backface_culling ? 1 : 0, bg, camv, 0, ref.data(), 8, nullptr);
    orc_ploc_free(ploc);
    return ref;
}

// Renders scene with the GPU Whitted renderer and with the oracle's restatement; true when
// every value is bit-identical.
bool WhittedMatchesOracle(const GpuBvh& bvh, const RenderConfig& rc, const Camera& cam, std::span<const Triangle> tris,
                          std::span<const Material> mats, double* sum) {
    const std::vector<float> img = GpuWhittedRenderer(bvh).Render(rc, cam, tris, mats);
    const std::vector<float> ref = OracleWhitted(rc, cam, tris, mats);
    size_t diff = 0;
    *sum = 0.0;
    for (size_t i = 0; i < ref.size(); i++) {
        diff += std::memcmp(&img[i], &ref[i], 4) != 0;
        *sum += img[i];
    }
    return diff == 0;
}

void GpuChecks() {
    auto ctx = std::make_shared<Context>(0);
    const std::vector<Triangle> tris = MakeScene(20000, 7);
    std::vector<float> xyz;
    for (const Triangle& t : tris)
        for (const Vec3* v : {&t.v0.pos, &t.v1.pos, &t.v2.pos}) xyz.insert(xyz.end(), {v->x, v->y, v->z});
    const std::vector<Ray> rays = CameraRays(160, 120);
    std::vector<dpl_ray> rr;
    for (const Ray& r : rays) rr.push_back({r.origin.x, r.origin.y, r.origin.z, kEpsilon, r.dir.x, r.dir.y, r.dir.z, r.t_max});
    std::vector<dpl_hit> brute(rays.size());
    orc_trace_brute(xyz.data(), int64_t(tris.size()), rr.data(), int64_t(rr.size()), 0, brute.data(), 8);

    for (const char* name : {"ploc", "ploc_emc", "ploc_sobb", "ploc_emc_sobb"}) {
        AccelerationStructureConfig cfg;
        cfg.acceleration_structure_type = ParseAccelerationStructureType(name);
        cfg.nn_search_radius = 16;
        std::unique_ptr<AccelerationStructure> as = CreateGpuAccelerationStructure(cfg, ctx);
        Stats stats;
        as->Build(cfg, stats, tris);
        Expect(stats.accel_stats.build_time > 0.0 && stats.accel_stats.node_count ==
                   stats.accel_stats.inner_node_count + stats.accel_stats.leaf_node_count,
               std::string(name) + ": build stats");
        auto* gpu = dynamic_cast<GpuBvh*>(as.get());
        std::vector<RayHit> hits(rays.size());
        const int64_t nhit = gpu->IntersectBatch(stats, rays, hits, false);
        Expect(stats.accel_stats.query_count == int(rays.size()) && stats.accel_stats.traversal_count > 0,
               std::string(name) + ": query stats");
        int64_t mismatch = 0, expect_hits = 0;
        for (size_t i = 0; i < rays.size(); i++) {
            expect_hits += brute[i].prim >= 0;
            if (hits[i].triangle_id != brute[i].prim) mismatch++;
            if (hits[i].triangle_id >= 0) {
                const Triangle& t = tris[size_t(hits[i].triangle_id)];
                if (hits[i].t != brute[i].t || hits[i].material_id != t.material_id ||
                    std::fabs(hits[i].b0 + hits[i].b1 + hits[i].b2 - 1.0f) > 1e-5f)
                    mismatch++;
            }
        }
        Expect(mismatch == 0 && nhit == expect_hits, std::string(name) + ": batched hits equal brute force");
        // the per-ray slow path agrees with the batch
        int64_t bad = 0;
        for (size_t i = 0; i < rays.size(); i += 97) {
            RayHit h;
            const bool hit = as->Intersect(stats, rays[i], h, false);
            if (hit != (hits[i].triangle_id >= 0) || (hit && (h.triangle_id != hits[i].triangle_id || h.t != hits[i].t)))
                bad++;
            if (as->IntersectAny(stats, rays[i], false) != hit) bad++;
        }
        std::vector<uint8_t> occ(rays.size());
        gpu->IntersectAnyBatch(stats, rays, occ, false);
        for (size_t i = 0; i < rays.size(); i++) bad += occ[i] != (hits[i].triangle_id >= 0);
        Expect(bad == 0, std::string(name) + ": Intersect / IntersectAny agree with the batch");
        // the wide form (4 and 8 children) gives the same closest hits; a rebuild drops it
        for (int width : {4, 8}) {
            gpu->Widen(width);
            std::vector<RayHit> wh(rays.size());
            const int64_t wn = gpu->IntersectBatch(stats, rays, wh, false);
            int64_t wbad = wn != nhit;
            for (size_t i = 0; i < rays.size(); i++)
                wbad += wh[i].triangle_id != hits[i].triangle_id || (wh[i].triangle_id >= 0 && wh[i].t != hits[i].t);
            gpu->IntersectAnyBatch(stats, rays, occ, false);
            for (size_t i = 0; i < rays.size(); i++) wbad += occ[i] != (hits[i].triangle_id >= 0);
            Expect(wbad == 0 && gpu->wide_width() == width,
                   std::string(name) + ": wide" + std::to_string(width) + " hits equal the two-child trace");
        }
        as->Build(cfg, stats, tris);
        Expect(gpu->wide_width() == 0, std::string(name) + ": a rebuild drops the wide form");
        std::printf("%s: %lld hits of %zu rays, build %.3f ms, %d nodes\n", name, (long long)nhit, rays.size(),
                    stats.accel_stats.build_time, stats.accel_stats.node_count);
        std::fflush(stdout);
    }

    // path tracing and its adjoint through the same BVH
    AccelerationStructureConfig cfg;
    cfg.acceleration_structure_type = AccelerationStructureType::kPlocEmc;
    GpuBvh bvh(ctx);
    Stats stats;
    bvh.Build(cfg, stats, tris);
    const int M = 7;
    std::vector<float> albedo(3 * M, 0.5f), emission(3 * M, 0.0f);
    emission[3 * 1] = emission[3 * 1 + 1] = emission[3 * 1 + 2] = 5.0f;  // the ceiling glows
    GpuPathTracer pt(bvh, tris, M);
    PathTraceConfig pc{64, 48, 4, 4, 3};
    Camera cam{{0.5f, 0.5f, 0.05f}, {0, 0, 1}, {0, 1, 0}, 1.2f, 100.0f};
    const std::vector<float> img = pt.Render(pc, cam, albedo, emission);
    const std::vector<float> img2 = pt.Render(pc, cam, albedo, emission);
    double sum = 0.0;
    bool finite = true;
    for (float v : img) {
        sum += v;
        finite &= std::isfinite(v);
    }
    Expect(finite && sum > 0.0 && img == img2, "path tracer: finite, lit and deterministic");
    // L = sum(image): dL/dimage = 1; every albedo gradient is non-negative, the emitter's is positive
    std::vector<float> ones(img.size(), 1.0f), ga(3 * M), ge(3 * M);
    pt.Adjoint(pc, cam, albedo, emission, ones, ga, ge);
    bool nonneg = true;
    for (float g : ga) nonneg &= g >= 0.0f;
    Expect(nonneg && ge[3] > 0.0f, "path tracer adjoint: gradient signs");
    // emission enters linearly: L(emission) = sum(image) = dL/de . e for the one emitter
    const double lin = double(ge[3]) * 5.0 + double(ge[4]) * 5.0 + double(ge[5]) * 5.0;
    Expect(std::fabs(lin - sum) <= 1e-3 * std::fabs(sum), "path tracer adjoint: linearity in emission");
    std::printf("path tracer: image sum %.4f, dL/de %.4f %.4f %.4f\n", sum, ge[3], ge[4], ge[5]);

    // Whitted renderer (the reference's CpuRaytracer) against the oracle's restatement
    std::vector<Material> wm(7);
    for (int m = 0; m < 7; m++) wm[m].diffuse = {0.7f, 0.6f - 0.05f * m, 0.5f};
    wm[1].diffuse = {};
    wm[1].emission = {6.0f, 6.0f, 5.0f};  // the ceiling
    wm[2].specular = {0.5f, 0.5f, 0.5f}, wm[2].shininess = 32.0f;
    wm[6].specular = {0.2f, 0.2f, 0.2f}, wm[6].transmittance = {0.6f, 0.6f, 0.6f};
    wm[6].ior = 1.5f, wm[6].r_ior = 1.0f / 1.5f, wm[6].shininess = 16.0f;
    RenderConfig rc;
    rc.width = 96, rc.height = 72, rc.max_depth = 3, rc.area_light_sample_cnt = 2;
    rc.background_color = {0.05f, 0.1f, 0.2f};
    const std::vector<float> wimg = GpuWhittedRenderer(bvh).Render(rc, cam, tris, wm);
    std::vector<float> mrows(16 * 7, 0.0f), vn(tris.size() * 9, 0.0f);
    std::vector<int32_t> mid(tris.size());
    for (int m = 0; m < 7; m++) {
        const Vec3 f[4] = {wm[m].diffuse, wm[m].specular, wm[m].transmittance, wm[m].emission};
        for (int k = 0; k < 4; k++) mrows[16 * m + 3 * k] = f[k].x, mrows[16 * m + 3 * k + 1] = f[k].y, mrows[16 * m + 3 * k + 2] = f[k].z;
        mrows[16 * m + 12] = wm[m].shininess, mrows[16 * m + 13] = wm[m].ior, mrows[16 * m + 14] = wm[m].r_ior;
    }
    for (size_t i = 0; i < tris.size(); i++) {
        mid[i] = tris[i].material_id;
        if (tris[i].has_vertex_normals)
            for (int k = 0; k < 3; k++) {
                const Vec3& nv = k == 0 ? tris[i].v0.normal : (k == 1 ? tris[i].v1.normal : tris[i].v2.normal);
                vn[9 * i + 3 * k] = nv.x, vn[9 * i + 3 * k + 1] = nv.y, vn[9 * i + 3 * k + 2] = nv.z;
            }
    }
    void* ploc = orc_ploc_build(xyz.data(), int64_t(tris.size()), 0, 16, 4, 0, 32, nullptr);
    std::vector<float> ref(wimg.size());
    const float bg[3] = {0.05f, 0.1f, 0.2f};
    const float camv[11] = {cam.pos.x, cam.pos.y, cam.pos.z, cam.dir.x, cam.dir.y, cam.dir.z,
                            cam.up.x,  cam.up.y,  cam.up.z,  cam.fov,   cam.far};
    orc_whitted(ploc, xyz.data(), vn.data(), int64_t(tris.size()), mid.data(), mrows.data(), rc.width, rc.height,
                rc.max_depth, rc.area_light_sample_cnt, 0, bg, camv, 0, ref.data(), 8, nullptr);
    orc_ploc_free(ploc);
    size_t diff = 0;
    double wsum = 0.0;
    for (size_t i = 0; i < ref.size(); i++) {
        diff += std::memcmp(&wimg[i], &ref[i], 4) != 0;
        wsum += wimg[i];
    }
    Expect(diff == 0 && wsum > 0.0, "whitted renderer: bit-identical to the oracle");
    std::printf("whitted: %zu of %zu values differ, image sum %.4f\n", diff, ref.size(), wsum);

    if (!g_scene_dir.empty()) {
        // load -> build through the factory -> render, the reference's renderer path with its
        // sample configuration (max depth 8, 8 area-light samples)
        SceneLoadConfig sc;
        sc.dirpath = g_scene_dir;
        sc.name = "CornellBox-Sphere";
        const auto scene = CreateSceneLoader(sc)->Load(sc);
        Expect(scene.has_value(), "scene loader (GPU path): " + ObjSceneLoader::LastError());
        if (scene) {
            AccelerationStructureConfig ac;
            ac.acceleration_structure_type = AccelerationStructureType::kPlocEmcSobb;
            auto as = CreateGpuAccelerationStructure(ac, ctx);
            Stats st;
            as->Build(ac, st, scene->Triangles());
            RenderConfig src;
            src.width = 96, src.height = 96, src.max_depth = 8, src.area_light_sample_cnt = 8;
            double ssum = 0.0;
            const bool same = WhittedMatchesOracle(static_cast<const GpuBvh&>(*as), src, scene->GetCamera(),
                                                   scene->Triangles(), scene->Materials(), &ssum);
            Expect(same && ssum > 0.0, "loaded CornellBox-Sphere: Whitted bit-identical to the oracle");
            // the renderer's counters, as the reference's renderer keeps them in Stats
            Stats rs;
            const std::vector<float> img2 = GpuWhittedRenderer(static_cast<const GpuBvh&>(*as))
                                                .Render(src, scene->GetCamera(), scene->Triangles(),
                                                        scene->Materials(), &rs);
            const RaytracingStats& r = rs.rt_stats;
            Expect(r.primary_ray_count == 96 * 96 && r.secondary_ray_count > 0 && r.shadow_ray_count > 0 &&
                       r.raytracing_time > 0.0,
                   "whitted stats: ray counters");
            Expect(rs.accel_stats.query_count == r.primary_ray_count + r.secondary_ray_count + r.shadow_ray_count &&
                       rs.accel_stats.traversal_count > rs.accel_stats.query_count,
                   "whitted stats: every traced ray is one query");
            std::printf("whitted stats: %d primary, %d secondary, %d shadow rays, %.3f ms\n", r.primary_ray_count,
                        r.secondary_ray_count, r.shadow_ray_count, r.raytracing_time * 1e3);
            std::printf("loaded scene: %zu triangles, %d nodes, whitted image sum %.4f\n", scene->Triangles().size(),
                        st.accel_stats.node_count, ssum);
        }
        // the reference's Renderer interface: load -> GpuRenderer::StartRender(RenderConfig,
        // AccelerationStructureConfig, Scene, Framebuffer, Stats), as main() drives a renderer
        for (const char* name : {"CornellBox-Original", "CornellBox-Sphere"}) {
            sc.name = name;
            const auto sc2 = CreateSceneLoader(sc)->Load(sc);
            Expect(sc2.has_value(), std::string("scene loader (renderer): ") + ObjSceneLoader::LastError());
            if (!sc2) continue;
            for (AccelerationStructureType type : {AccelerationStructureType::kPlocEmcSobb, AccelerationStructureType::kSbvh}) {
                RenderConfig rc2;
                rc2.renderer_type = RendererType::kGpu;
                rc2.width = 80, rc2.height = 64, rc2.max_depth = 8, rc2.area_light_sample_cnt = 8;
                rc2.background_color = Vec3{0.1f, 0.2f, 0.3f};
                AccelerationStructureConfig ac2;
                ac2.acceleration_structure_type = type;
                GpuRenderer renderer(ctx);
                Framebuffer fb;
                Stats rs;
                Expect(std::string(renderer.GetName()) == "GPU Raytracer" && renderer.GetProgress() == 0.0f,
                       "renderer: name and initial progress");
                const RenderResult res = renderer.StartRender(rc2, ac2, *sc2, fb, rs);
                Expect(res == RenderResult::kDone, std::string("renderer: StartRender done ") + renderer.LastError());
                const std::vector<float> ref = OracleWhitted(rc2, sc2->GetCamera(), sc2->Triangles(), sc2->Materials());
                size_t diff = 0;
                for (size_t i = 0; i < ref.size() && fb.GetData().size() == ref.size(); i++)
                    diff += std::memcmp(&fb.GetData()[i], &ref[i], 4) != 0;
                Expect(fb.GetWidth() == 80 && fb.GetHeight() == 64 && fb.GetData().size() == ref.size() && diff == 0,
                       std::string("renderer: ") + name + " framebuffer bit-identical to the oracle");
                const Vec3 p = fb.GetPixel(3, 5);
                Expect(p.x == ref[(5 * 80 + 3) * 3] && p.z == ref[(5 * 80 + 3) * 3 + 2], "renderer: GetPixel layout");
                const RaytracingStats& r = rs.rt_stats;
                // CornellBox-Original has no mirror or glass material: no secondary rays
                const bool mirrors = std::string(name) == "CornellBox-Sphere";
                Expect(r.primary_ray_count == 80 * 64 && (r.secondary_ray_count > 0) == mirrors &&
                           r.shadow_ray_count > 0 && r.raytracing_time > 0.0 && renderer.GetProgress() == 1.0f,
                       "renderer: ray counters and progress");
                Expect(rs.accel_stats.node_count > 0 && rs.accel_stats.build_time > 0.0 &&
                           rs.accel_stats.query_count == r.primary_ray_count + r.secondary_ray_count + r.shadow_ray_count,
                       "renderer: acceleration stats");
                std::printf("renderer %s: %d primary, %d secondary, %d shadow rays, %.3f ms, %zu values differ\n", name,
                            r.primary_ray_count, r.secondary_ray_count, r.shadow_ray_count, r.raytracing_time * 1e3,
                            diff);
            }
            GpuRenderer canceled(ctx);
            canceled.Cancel();
            Framebuffer fb;
            Stats rs;
            Expect(canceled.StartRender(RenderConfig{}, AccelerationStructureConfig{}, *sc2, fb, rs) ==
                       RenderResult::kCanceled && canceled.GetProgress() == 0.0f,
                   "renderer: canceled before the launch");
            canceled.Reset();
            RenderConfig bad;
            bad.width = 0;
            Expect(canceled.StartRender(bad, AccelerationStructureConfig{}, *sc2, fb, rs) == RenderResult::kError &&
                       !canceled.LastError().empty(),
                   "renderer: an invalid size is an error");
        }
    }
}

}  // namespace

int main(int argc, char** argv) {
    bool cpu_only = false;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--cpu") cpu_only = true;
        if (a == "--scenes" && i + 1 < argc) g_scene_dir = argv[++i];
        if (a == "--configs" && i + 1 < argc) g_config_dir = argv[++i];
    }
    CpuChecks();
    if (!cpu_only) {
        try {
            GpuChecks();
        } catch (const std::exception& e) {
            std::printf("FAILED: exception %s\n", e.what());
            g_failures++;
        }
    }
    std::printf("host_api_test: %s\n", g_failures ? "FAILED" : "OK");
    return g_failures ? 1 : 0;
}
