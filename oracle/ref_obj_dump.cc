// Test infrastructure only: the reference's scene loading, compiled from the reference's own
// vendored OBJ parser (third-party/tinyobjloader/tiny_obj_loader.h, used through
// tinyobj::ObjReader with its default ObjReaderConfig as src/io/scene/obj_scene_loader.cc:20-35
// does). The conversion of tinyobj's shapes and materials into triangles, area lights and
// materials restates ObjSceneLoader::LoadObj (obj_scene_loader.cc:36-125), and the .meta
// reader restates ObjSceneLoader::LoadMetadata (obj_scene_loader.cc:128-186); the reference's
// own file cannot be compiled here (its logger needs <format>, which this toolchain lacks).
//
// Usage: ref_obj_dump <dir> <name>    reads <dir>/<name>.meta and <dir>/<name>.obj
//        ref_obj_dump --obj <file.obj> reads one OBJ (and its MTL) only
// Prints one JSON object; every float is printed as its IEEE-754 bit pattern (uint32).
#define TINYOBJLOADER_IMPLEMENTATION
#include "tiny_obj_loader.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <string>
#include <vector>

namespace {

uint32_t Bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

struct V3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};
V3 Sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 Cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float Length(V3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }

struct Tri {
    V3 p[3], n[3];
    bool has_n = false;
    int mat = -1;
};

struct Out {
    std::vector<Tri> tris;
    std::vector<tinyobj::material_t> mats;
    std::vector<std::pair<int, float>> area_lights;  // triangle id, surface area
    std::vector<float> point_lights;                 // pos3 color3 power per light
    float cam[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
};

// obj_scene_loader.cc:128-186
bool Meta(const std::filesystem::path& path, Out& o) {
    std::ifstream in(path);
    if (!in) return false;
    std::string tok;
    float* c = o.cam;
    while (in >> tok) {
        if (tok == "cam_pos") {
            if (!(in >> c[0] >> c[1] >> c[2])) return false;
        } else if (tok == "cam_up") {
            if (!(in >> c[6] >> c[7] >> c[8])) return false;
        } else if (tok == "cam_dir") {
            if (!(in >> c[3] >> c[4] >> c[5])) return false;
        } else if (tok == "cam_fov") {
            if (!(in >> c[9])) return false;
        } else if (tok == "cam_far") {
            if (!(in >> c[10])) return false;
        } else if (tok == "point_light") {
            float v[7];
            if (!(in >> v[0] >> v[1] >> v[2])) return false;
            if (!(in >> v[3] >> v[4] >> v[5])) return false;
            if (!(in >> v[6])) return false;
            o.point_lights.insert(o.point_lights.end(), v, v + 7);
        } else {
            return false;
        }
    }
    return true;
}

// obj_scene_loader.cc:20-125
bool Obj(const std::filesystem::path& path, Out& o) {
    tinyobj::ObjReaderConfig cfg;
    cfg.mtl_search_path = path.parent_path().string();
    tinyobj::ObjReader reader;
    if (!reader.ParseFromFile(path.string(), cfg)) return false;
    const auto& a = reader.GetAttrib();
    for (const auto& shape : reader.GetShapes()) {
        for (size_t f = 0; f < shape.mesh.num_face_vertices.size(); f++) {
            if (shape.mesh.num_face_vertices[f] != 3) return false;
            Tri t;
            tinyobj::index_t idx[3] = {shape.mesh.indices[3 * f], shape.mesh.indices[3 * f + 1],
                                       shape.mesh.indices[3 * f + 2]};
            for (int k = 0; k < 3; k++) {
                const size_t vi = size_t(idx[k].vertex_index);
                t.p[k] = {a.vertices[3 * vi], a.vertices[3 * vi + 1], a.vertices[3 * vi + 2]};
            }
            if (idx[0].normal_index >= 0 && idx[1].normal_index >= 0 && idx[2].normal_index >= 0) {
                for (int k = 0; k < 3; k++) {
                    const size_t ni = size_t(idx[k].normal_index);
                    t.n[k] = {a.normals[3 * ni], a.normals[3 * ni + 1], a.normals[3 * ni + 2]};
                }
                t.has_n = true;
            }
            t.mat = shape.mesh.material_ids[f];
            o.tris.push_back(t);
            if (t.mat >= 0) {
                const auto& m = reader.GetMaterials()[size_t(t.mat)];
                if (m.emission[0] != 0.0f || m.emission[1] != 0.0f || m.emission[2] != 0.0f) {
                    const float area = 0.5 * Length(Cross(Sub(t.p[1], t.p[0]), Sub(t.p[2], t.p[0])));
                    o.area_lights.push_back({int(o.tris.size()) - 1, area});
                }
            }
        }
    }
    o.mats = reader.GetMaterials();
    return true;
}

void PrintBits(const float* v, int n) {
    std::printf("[");
    for (int i = 0; i < n; i++) std::printf(i ? ",%u" : "%u", Bits(v[i]));
    std::printf("]");
}

void Print(const Out& o, bool ok) {
    std::printf("{\"ok\":%s,\"camera\":", ok ? "true" : "false");
    PrintBits(o.cam, 11);
    std::printf(",\"point_lights\":");
    PrintBits(o.point_lights.data(), int(o.point_lights.size()));
    std::printf(",\"triangles\":[");
    for (size_t i = 0; i < o.tris.size(); i++) {
        const Tri& t = o.tris[i];
        float v[18];
        for (int k = 0; k < 3; k++) {
            v[3 * k] = t.p[k].x, v[3 * k + 1] = t.p[k].y, v[3 * k + 2] = t.p[k].z;
            v[9 + 3 * k] = t.n[k].x, v[10 + 3 * k] = t.n[k].y, v[11 + 3 * k] = t.n[k].z;
        }
        std::printf(i ? ",{\"v\":" : "{\"v\":");
        PrintBits(v, 18);
        std::printf(",\"has_n\":%d,\"mat\":%d}", int(t.has_n), t.mat);
    }
    std::printf("],\"materials\":[");
    for (size_t i = 0; i < o.mats.size(); i++) {
        const auto& m = o.mats[i];
        // ambient3 diffuse3 specular3 transmittance3 emission3 shininess ior r_ior dissolve
        float v[19];
        for (int k = 0; k < 3; k++) {
            v[k] = m.ambient[k], v[3 + k] = m.diffuse[k], v[6 + k] = m.specular[k];
            v[9 + k] = m.transmittance[k], v[12 + k] = m.emission[k];
        }
        v[15] = m.shininess, v[16] = m.ior, v[17] = 1.0f / m.ior, v[18] = m.dissolve;
        std::printf(i ? ",{\"name\":\"%s\",\"v\":" : "{\"name\":\"%s\",\"v\":", m.name.c_str());
        PrintBits(v, 19);
        std::printf("}");
    }
    std::printf("],\"area_lights\":[");
    for (size_t i = 0; i < o.area_lights.size(); i++)
        std::printf(i ? ",[%d,%u]" : "[%d,%u]", o.area_lights[i].first, Bits(o.area_lights[i].second));
    std::printf("]}\n");
}

}  // namespace

int main(int argc, char** argv) {
    Out o;
    bool ok = false;
    if (argc == 3 && std::strcmp(argv[1], "--obj") == 0) {
        ok = Obj(argv[2], o);
    } else if (argc == 3) {
        const std::filesystem::path base = std::filesystem::path(argv[1]) / argv[2];
        ok = Meta(std::filesystem::path(base.string() + ".meta"), o) &&
             Obj(std::filesystem::path(base.string() + ".obj"), o);
    } else {
        std::fprintf(stderr, "usage: ref_obj_dump <dir> <name> | --obj <file.obj>\n");
        return 2;
    }
    Print(o, ok);
    return 0;
}
