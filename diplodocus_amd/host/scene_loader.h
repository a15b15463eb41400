// Scene loading for the engine's host interface: the reference's Scene container and its OBJ
// scene loader, so that a renderer reads the same files into the same triangles, materials,
// lights and camera.
//   * Scene, PointLight, AreaLight: src/scene/scene.h:11-38, src/scene/light.h:9-21;
//   * SceneLoadFormat / SceneLoadConfig: src/config/scene_load_config.h:11-24;
//   * SceneLoader / ObjSceneLoader / CreateSceneLoader: src/io/scene/scene_loader.h:9-13,
//     obj_scene_loader.h:7-15, scene_loader_factory.cc:11-19.
// Load(config) reads <dirpath>/<name>.meta (camera, point lights) and then <dirpath>/<name>.obj
// with the material libraries it names, as obj_scene_loader.cc:188-204 does. The OBJ reading
// gives the triangles the reference's tinyobjloader-based loader gives for the same file:
// polygons split the way tinyobjloader splits them (quads along the shorter diagonal, larger
// polygons by its ear clipping), numbers rounded the same way, faces before any `usemtl` or
// with an unknown material get material_id -1, and every emissive triangle becomes an area
// light in triangle order. Where the reference's behaviour is undefined or partial, loading
// fails instead: a face index beyond the vertices or normals defined before it, a malformed
// number in an OBJ or MTL file, a .meta token that is not a keyword or a number where one is
// due. Lines (`l`), points (`p`), texture maps other than map_Kd's default, and the MTL keys a
// Material does not hold are read past. Failures return std::nullopt / false with the reason
// in ObjSceneLoader::LastError() (per thread).
#pragma once

#include <filesystem>
#include <memory>
#include <optional>
#include <span>
#include <string>
#include <string_view>
#include <vector>

#include "dpl_host.h"

namespace diplodocus_amd {

struct PointLight {
    Vec3 pos;
    Vec3 color;
    float power = 0.0f;
};

struct AreaLight {
    Vec3 color;
    int triangle_id = -1;
    float surface_area = 0.0f;
    float power = 0.0f;  // not set by the OBJ loader (as in the reference)
};

// Unit normal of the triangle's plane, or zero for a degenerate triangle
// (RecomputeTriangleGeometricNormal, src/scene/triangle.h:22-26).
void RecomputeTriangleGeometricNormal(Triangle& t);
// Half the length of the edge cross product (CalculateTriangleSurfaceArea, triangle.h:28-34).
float CalculateTriangleSurfaceArea(const Triangle& t);

class Scene {
   public:
    void ReserveTriangles(size_t n) { triangles_.reserve(n); }
    void ReserveMaterials(size_t n) { materials_.reserve(n); }
    void AddTriangle(Triangle t);  // recomputes geom_normal
    void AddMaterial(Material m) { materials_.push_back(std::move(m)); }
    void AddPointLight(PointLight l) { point_lights_.push_back(l); }
    void AddAreaLight(AreaLight l) { area_lights_.push_back(l); }

    Camera& GetCamera() { return camera_; }
    const Camera& GetCamera() const { return camera_; }
    std::span<const Triangle> Triangles() const { return triangles_; }
    std::span<const Material> Materials() const { return materials_; }
    std::span<const PointLight> PointLights() const { return point_lights_; }
    std::span<const AreaLight> AreaLights() const { return area_lights_; }

   private:
    Camera camera_;
    std::vector<Triangle> triangles_;
    std::vector<Material> materials_;
    std::vector<PointLight> point_lights_;
    std::vector<AreaLight> area_lights_;
};

enum class SceneLoadFormat { kObj };

// "obj"; throws std::runtime_error on anything else.
SceneLoadFormat ParseSceneLoadFormat(std::string_view sv);

struct SceneLoadConfig {
    SceneLoadFormat scene_load_format = SceneLoadFormat::kObj;
    std::filesystem::path dirpath = "res";
    std::string name = "CornellBox-Sphere";
    bool triangulate = true;  // polygons are always split, as in the reference
};

class SceneLoader {
   public:
    virtual ~SceneLoader() = default;
    virtual std::optional<Scene> Load(const SceneLoadConfig& config) const = 0;
};

class ObjSceneLoader final : public SceneLoader {
   public:
    std::optional<Scene> Load(const SceneLoadConfig& config) const override;
    // Appends the OBJ's triangles, area lights and materials to scene.
    static bool LoadObj(const std::filesystem::path& obj_file_path, Scene& scene);
    // Reads the camera and point lights of a .meta file into scene.
    static bool LoadMetadata(const std::filesystem::path& metadata_file_path, Scene& scene);
    static const std::string& LastError();
};

// The factory (scene_loader_factory.cc:11-19); throws std::invalid_argument for an unknown
// format.
std::unique_ptr<SceneLoader> CreateSceneLoader(const SceneLoadConfig& config);

}  // namespace diplodocus_amd
