// OBJ / MTL / .meta scene loading (scene_loader.h) and its C ABI (include/dpl/dpl_scene.h).
//
// The text is read whole and walked line by line; each recognised line is parsed in place.
// Faces are kept as runs of corners (position, texcoord, normal indices) with the material
// current when the face was read, and split into triangles once the whole file is read.
// Numbers, indices and polygon splitting follow the rules of the reference's vendored
// tinyobjloader (third-party/tinyobjloader/tiny_obj_loader.h) so that a file yields the same
// triangles bit for bit:
//   * a number is the longest decimal prefix of its token ([+-] digits [. digits]
//     [e [+-] digits]), correctly rounded to double and then to float; "nan" is 0 and
//     "inf"/"infinity" the largest double; a non-zero value whose double is 0 or infinite is
//     an error (tryParseDouble, :5964-5996; tryParseNanInf, :5906-5948);
//   * an index is [+-]digits (0 when absent or outside int32), 1-based, negative counting
//     back from the last element read; a zero position index is an error, a zero texcoord or
//     normal index means none (sr_parseIntNoSkip :6752-6777, fixIndex :5834-5867);
//   * a quad is split along its shorter diagonal and a larger polygon by ear clipping in the
//     plane picked from its first non-degenerate corner, in float arithmetic
//     (exportGroupsToShape :7080-7600).
#include "scene_loader.h"

#include <algorithm>
#include <array>
#include <charconv>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <locale>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>

#include "dpl/dpl_scene.h"

namespace diplodocus_amd {

namespace {

thread_local std::string t_error;

bool Fail(const std::string& what) {
    t_error = what;
    return false;
}

bool ReadFile(const std::filesystem::path& path, std::string& out) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    std::ostringstream ss;
    ss << in.rdbuf();
    out = ss.str();
    return true;
}

// A line's text; `p` advances as it is parsed.
struct Line {
    const char* p;
    const char* e;

    bool Blank(char c) const { return c == ' ' || c == '\t'; }
    void SkipBlank() {
        while (p < e && Blank(*p)) ++p;
    }
    bool AtEnd() const { return p >= e; }
    std::string_view Token() {
        SkipBlank();
        const char* s = p;
        while (p < e && !Blank(*p)) ++p;
        return {s, size_t(p - s)};
    }
    // True (and consumes it) when the line starts with kw followed by a blank.
    bool Keyword(std::string_view kw) {
        if (size_t(e - p) <= kw.size() || std::memcmp(p, kw.data(), kw.size()) != 0 || !Blank(p[kw.size()]))
            return false;
        p += kw.size();
        return true;
    }
};

// Calls f(Line) for each line: "\n", "\r\n" and a lone "\r" end a line, a NUL ends its
// content, and a leading UTF-8 byte order mark is skipped.
template <class F>
bool ForEachLine(const std::string& text, F&& f) {
    const char* p = text.data();
    const char* end = p + text.size();
    if (text.size() >= 3 && std::memcmp(p, "\xEF\xBB\xBF", 3) == 0) p += 3;
    while (p < end) {
        const char* q = p;
        while (q < end && *q != '\n' && *q != '\r') ++q;
        const char* nul = static_cast<const char*>(std::memchr(p, '\0', size_t(q - p)));
        if (!f(Line{p, nul ? nul : q})) return false;
        p = q;
        if (p < end && *p == '\r') ++p;
        if (p < end && *p == '\n') ++p;
    }
    return true;
}

bool StartsWithNoCase(std::string_view s, std::string_view prefix) {
    if (s.size() < prefix.size()) return false;
    for (size_t i = 0; i < prefix.size(); i++) {
        char c = s[i];
        if (c >= 'A' && c <= 'Z') c = char(c + 32);
        if (c != prefix[i]) return false;
    }
    return true;
}

// A number token to float; false for a malformed or out-of-range number.
bool ToReal(std::string_view tok, float* out) {
    const bool neg = !tok.empty() && tok[0] == '-';
    std::string_view body = (!tok.empty() && (tok[0] == '+' || tok[0] == '-')) ? tok.substr(1) : tok;
    if (StartsWithNoCase(body, "nan")) {
        *out = 0.0f;
        return true;
    }
    if (StartsWithNoCase(body, "inf")) {
        const double big = std::numeric_limits<double>::max();
        *out = static_cast<float>(neg ? -big : big);
        return true;
    }
    // the decimal prefix: digits [. digits] or . digits, then an exponent only when it has
    // digits
    size_t i = 0, int_digits = 0, frac_digits = 0;
    while (i < body.size() && body[i] >= '0' && body[i] <= '9') i++, int_digits++;
    if (i < body.size() && body[i] == '.') {
        size_t j = i + 1;
        while (j < body.size() && body[j] >= '0' && body[j] <= '9') j++, frac_digits++;
        if (int_digits + frac_digits > 0) i = j;
    }
    if (int_digits + frac_digits == 0) return false;
    if (i < body.size() && (body[i] == 'e' || body[i] == 'E')) {
        size_t j = i + 1;
        if (j < body.size() && (body[j] == '+' || body[j] == '-')) j++;
        const size_t d0 = j;
        while (j < body.size() && body[j] >= '0' && body[j] <= '9') j++;
        if (j > d0) i = j;
    }
    double v = 0.0;
    const auto r = std::from_chars(body.data(), body.data() + i, v, std::chars_format::general);
    if (r.ec != std::errc() || r.ptr != body.data() + i) return false;  // out of double's range
    *out = static_cast<float>(neg ? -v : v);
    return true;
}

// sr_parseReal with a default of 0: an empty token reads as 0, a bad one is an error.
bool ReadReal(Line& ln, float* out) {
    const std::string_view tok = ln.Token();
    if (tok.empty()) {
        *out = 0.0f;
        return true;
    }
    return ToReal(tok, out);
}

bool ReadReals(Line& ln, float* out, int n) {
    for (int i = 0; i < n; i++)
        if (!ReadReal(ln, out + i)) return false;
    return true;
}

// [+-]digits at the cursor; 0 when there are none, more than 63 characters or the value is
// outside int32.
int ReadIndexInt(Line& ln) {
    const char* s = ln.p;
    const char* q = s;
    if (q < ln.e && (*q == '+' || *q == '-')) ++q;
    while (q < ln.e && *q >= '0' && *q <= '9') ++q;
    ln.p = q;
    const size_t len = size_t(q - s);
    if (len == 0 || len > 63) return 0;
    long long v = 0;
    const char* b = (*s == '+') ? s + 1 : s;
    const auto r = std::from_chars(b, q, v);
    if (r.ec != std::errc() || r.ptr != q || v > std::numeric_limits<int>::max() ||
        v < std::numeric_limits<int>::min())
        return 0;
    return int(v);
}

// 1-based or negative-relative index to 0-based; -1 for an allowed zero; false for an error.
bool FixIndex(int idx, int count, bool allow_zero, int* out) {
    if (idx > 0) {
        *out = idx - 1;
        return true;
    }
    if (idx == 0) {
        *out = -1;
        return allow_zero;
    }
    *out = count + idx;
    return *out >= 0;
}

void SkipIndexRest(Line& ln) {
    while (ln.p < ln.e && *ln.p != '/' && *ln.p != ' ' && *ln.p != '\t') ++ln.p;
}

struct Corner {
    int v = -1, vt = -1, vn = -1;
};

struct FaceRun {
    size_t first;
    int count;
    int material;
};

struct MaterialTable {
    std::vector<Material> list;
    std::map<std::string, int> by_name;  // the first definition of a name wins
};

bool ReadMtl(const std::filesystem::path& path, MaterialTable& table) {
    std::string text;
    if (!ReadFile(path, text)) return false;
    Material m;
    bool has_d = false, has_kd = false, ok = true;
    auto flush = [&] {
        if (m.name.empty()) return;
        table.by_name.emplace(m.name, int(table.list.size()));
        table.list.push_back(m);
    };
    auto vec3 = [&](Line& ln, Vec3& v) {
        float f[3];
        if (!ReadReals(ln, f, 3)) return false;
        v = {f[0], f[1], f[2]};
        return true;
    };
    ForEachLine(text, [&](Line ln) {
        ln.SkipBlank();
        if (ln.AtEnd() || *ln.p == '#') return true;
        if (ln.Keyword("newmtl")) {
            flush();
            m = Material{};
            m.shininess = 1.0f;
            has_d = has_kd = false;
            m.name = std::string(ln.Token());
            return true;
        }
        bool good = true;
        if (ln.Keyword("Ka")) {
            good = vec3(ln, m.ambient);
        } else if (ln.Keyword("Kd")) {
            good = vec3(ln, m.diffuse);
            has_kd = true;
        } else if (ln.Keyword("Ks")) {
            good = vec3(ln, m.specular);
        } else if (ln.Keyword("Kt") || ln.Keyword("Tf")) {
            good = vec3(ln, m.transmittance);
        } else if (ln.Keyword("Ke")) {
            good = vec3(ln, m.emission);
        } else if (ln.Keyword("Ni")) {
            good = ReadReal(ln, &m.ior);
        } else if (ln.Keyword("Ns")) {
            good = ReadReal(ln, &m.shininess);
        } else if (ln.Keyword("d")) {
            good = ReadReal(ln, &m.dissolve);
            has_d = true;
        } else if (ln.Keyword("Tr")) {
            if (!has_d) {
                float tr = 0.0f;
                good = ReadReal(ln, &tr);
                m.dissolve = 1.0f - tr;
            }
        } else if (ln.Keyword("map_Kd")) {
            if (!has_kd) m.diffuse = {0.6f, 0.6f, 0.6f};
        }
        if (!good) ok = Fail(path.string() + ": malformed number in material library");
        return good;
    });
    if (!ok) return false;
    flush();
    return true;
}

// mtllib's file names: split at blanks (' '), where a backslash before a blank or a
// backslash makes it literal.
std::vector<std::string> SplitLibraryNames(std::string_view rest) {
    while (!rest.empty() && (rest.back() == ' ' || rest.back() == '\t')) rest.remove_suffix(1);
    std::vector<std::string> names;
    std::string cur;
    for (size_t i = 0; i < rest.size(); i++) {
        const char c = rest[i];
        if (c == '\\' && i + 1 < rest.size() && (rest[i + 1] == ' ' || rest[i + 1] == '\\')) {
            cur += rest[++i];
        } else if (c == ' ') {
            if (!cur.empty()) names.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    if (!cur.empty()) names.push_back(cur);
    return names;
}

// Crossing-number test of (tx, ty) against a triangle (the W. R. Franklin formulation).
bool InsideTriangle(const float* xs, const float* ys, float tx, float ty) {
    bool in = false;
    for (int i = 0, j = 2; i < 3; j = i++) {
        if ((ys[i] > ty) != (ys[j] > ty) && tx < (xs[j] - xs[i]) * (ty - ys[i]) / (ys[j] - ys[i]) + xs[i]) in = !in;
    }
    return in;
}

using Tri3 = std::array<Corner, 3>;

// Splits one polygon into triangles (appended to out).
void SplitPolygon(const Corner* c, int n, const std::vector<float>& pos, std::vector<Tri3>& out) {
    auto P = [&](const Corner& k, int axis) { return pos[3 * size_t(k.v) + size_t(axis)]; };
    if (n < 3) return;
    if (n == 3) {
        out.push_back({c[0], c[1], c[2]});
        return;
    }
    if (n == 4) {
        float d02[3], d13[3];
        for (int a = 0; a < 3; a++) d02[a] = P(c[2], a) - P(c[0], a), d13[a] = P(c[3], a) - P(c[1], a);
        const float s02 = d02[0] * d02[0] + d02[1] * d02[1] + d02[2] * d02[2];
        const float s13 = d13[0] * d13[0] + d13[1] * d13[1] + d13[2] * d13[2];
        if (s02 < s13) {
            out.push_back({c[0], c[1], c[2]});
            out.push_back({c[0], c[2], c[3]});
        } else {
            out.push_back({c[0], c[1], c[3]});
            out.push_back({c[1], c[2], c[3]});
        }
        return;
    }
    // projection plane: drop the axis the first non-degenerate corner's normal leans on most
    int ax = 1, ay = 2;
    const float eps = std::numeric_limits<float>::epsilon();
    for (int k = 0; k < n; k++) {
        const Corner& a = c[k % n];
        const Corner& b = c[(k + 1) % n];
        const Corner& d = c[(k + 2) % n];
        float e0[3], e1[3];
        for (int t = 0; t < 3; t++) e0[t] = P(b, t) - P(a, t), e1[t] = P(d, t) - P(b, t);
        const float nx = std::fabs(e0[1] * e1[2] - e0[2] * e1[1]);
        const float ny = std::fabs(e0[2] * e1[0] - e0[0] * e1[2]);
        const float nz = std::fabs(e0[0] * e1[1] - e0[1] * e1[0]);
        if (nx > eps || ny > eps || nz > eps) {
            if (!(nx > ny && nx > nz)) {
                ax = 0;
                if (nz > nx && nz > ny) ay = 1;
            }
            break;
        }
    }
    std::vector<Corner> ring(c, c + n);
    size_t guess = 0;
    size_t budget = size_t(n);  // attempts left without cutting an ear
    size_t last_size = ring.size();
    while (ring.size() > 3 && budget > 0) {
        const size_t m = ring.size();
        if (guess >= m) guess -= m;
        if (last_size != m) {
            last_size = m;
            budget = m;
        } else {
            budget--;
        }
        Tri3 t{ring[guess % m], ring[(guess + 1) % m], ring[(guess + 2) % m]};
        float xs[3], ys[3];
        for (int k = 0; k < 3; k++) xs[k] = P(t[k], ax), ys[k] = P(t[k], ay);
        const float turn = (xs[1] - xs[0]) * (ys[2] - ys[1]) - (ys[1] - ys[0]) * (xs[2] - xs[1]);
        const float sense = (xs[0] * ys[1] - ys[0] * xs[1]) * 0.5f;
        if (turn * sense < 0.0f) {  // a reflex corner
            guess++;
            continue;
        }
        bool blocked = false;
        for (size_t o = 3; o < m && !blocked; o++) {
            const Corner& q = ring[(guess + o) % m];
            blocked = InsideTriangle(xs, ys, P(q, ax), P(q, ay));
        }
        if (blocked) {
            guess++;
            continue;
        }
        out.push_back(t);
        ring.erase(ring.begin() + std::ptrdiff_t((guess + 1) % m));
    }
    if (ring.size() == 3) out.push_back({ring[0], ring[1], ring[2]});
}

struct ObjData {
    std::vector<float> v, vn;
    int vt_count = 0;
    std::vector<Corner> corners;
    std::vector<FaceRun> faces;
    MaterialTable materials;
};

bool ParseObj(const std::filesystem::path& path, ObjData& d) {
    std::string text;
    if (!ReadFile(path, text)) return Fail("cannot read " + path.string());
    const std::filesystem::path search = path.parent_path();
    std::set<std::string> libraries;
    int material = -1;
    size_t line_no = 0;
    auto bad = [&](const char* what) { return Fail(path.string() + ":" + std::to_string(line_no) + ": " + what); };

    auto read_corner = [&](Line& ln, Corner& k) {
        ln.SkipBlank();
        if (!FixIndex(ReadIndexInt(ln), int(d.v.size() / 3), false, &k.v)) return false;
        SkipIndexRest(ln);
        if (ln.AtEnd() || *ln.p != '/') return true;
        ++ln.p;
        if (!ln.AtEnd() && *ln.p == '/') {  // v//vn
            ++ln.p;
            if (!FixIndex(ReadIndexInt(ln), int(d.vn.size() / 3), true, &k.vn)) return false;
            SkipIndexRest(ln);
            return true;
        }
        if (!FixIndex(ReadIndexInt(ln), d.vt_count, true, &k.vt)) return false;  // v/vt[/vn]
        SkipIndexRest(ln);
        if (ln.AtEnd() || *ln.p != '/') return true;
        ++ln.p;
        if (!FixIndex(ReadIndexInt(ln), int(d.vn.size() / 3), true, &k.vn)) return false;
        SkipIndexRest(ln);
        return true;
    };

    return ForEachLine(text, [&](Line ln) {
        line_no++;
        ln.SkipBlank();
        if (ln.AtEnd() || *ln.p == '#') return true;
        if (ln.Keyword("v")) {
            float xyz[3];
            if (!ReadReals(ln, xyz, 3)) return bad("malformed vertex");
            d.v.insert(d.v.end(), xyz, xyz + 3);
        } else if (ln.Keyword("vn")) {
            float xyz[3];
            if (!ReadReals(ln, xyz, 3)) return bad("malformed normal");
            d.vn.insert(d.vn.end(), xyz, xyz + 3);
        } else if (ln.Keyword("vt")) {
            float uv[2];
            if (!ReadReals(ln, uv, 2)) return bad("malformed texture coordinate");
            d.vt_count++;
        } else if (ln.Keyword("f")) {
            FaceRun f{d.corners.size(), 0, material};
            for (;;) {
                ln.SkipBlank();
                if (ln.AtEnd() || *ln.p == '#') break;
                Corner k;
                if (!read_corner(ln, k)) return bad("invalid face index");
                if (size_t(k.v) >= d.v.size() / 3 || (k.vn >= 0 && size_t(k.vn) >= d.vn.size() / 3))
                    return bad("face index beyond the elements defined before it");
                d.corners.push_back(k);
                f.count++;
            }
            d.faces.push_back(f);
        } else if (ln.Keyword("usemtl")) {
            const auto it = d.materials.by_name.find(std::string(ln.Token()));
            material = it == d.materials.by_name.end() ? -1 : it->second;
        } else if (ln.Keyword("mtllib")) {
            // the names start one character after the keyword (tinyobjloader skips exactly one)
            for (const std::string& name : SplitLibraryNames(std::string_view(ln.p + 1, size_t(ln.e - ln.p - 1)))) {
                if (libraries.count(name)) continue;  // already read: try the next name
                const std::filesystem::path lib = search.empty() ? std::filesystem::path(name) : search / name;
                std::error_code ec;
                if (!std::filesystem::is_regular_file(lib, ec)) continue;
                if (!ReadMtl(lib, d.materials)) return false;
                libraries.insert(name);
                break;
            }
        }
        return true;
    });
}

}  // namespace

void RecomputeTriangleGeometricNormal(Triangle& t) {
    t.geom_normal = Normalize(Cross(t.v1.pos - t.v0.pos, t.v2.pos - t.v0.pos));
}

float CalculateTriangleSurfaceArea(const Triangle& t) {
    return 0.5f * Length(Cross(t.v1.pos - t.v0.pos, t.v2.pos - t.v0.pos));
}

void Scene::AddTriangle(Triangle t) {
    RecomputeTriangleGeometricNormal(t);
    triangles_.push_back(t);
}

SceneLoadFormat ParseSceneLoadFormat(std::string_view sv) {
    if (sv == "obj") return SceneLoadFormat::kObj;
    throw std::runtime_error("Invalid SceneLoadFormat: " + std::string(sv));
}

const std::string& ObjSceneLoader::LastError() { return t_error; }

bool ObjSceneLoader::LoadObj(const std::filesystem::path& obj_file_path, Scene& scene) {
    ObjData d;
    if (!ParseObj(obj_file_path, d)) return false;
    std::vector<Tri3> tris;
    std::vector<int> tri_material;
    for (const FaceRun& f : d.faces) {
        SplitPolygon(d.corners.data() + f.first, f.count, d.v, tris);
        tri_material.resize(tris.size(), f.material);
    }
    scene.ReserveTriangles(scene.Triangles().size() + tris.size());
    auto vec = [](const std::vector<float>& a, int i) { return Vec3{a[3 * size_t(i)], a[3 * size_t(i) + 1], a[3 * size_t(i) + 2]}; };
    for (size_t i = 0; i < tris.size(); i++) {
        const Tri3& c = tris[i];
        Triangle t;
        Vertex* vs[3] = {&t.v0, &t.v1, &t.v2};
        for (int k = 0; k < 3; k++) vs[k]->pos = vec(d.v, c[k].v);
        t.has_vertex_normals = c[0].vn >= 0 && c[1].vn >= 0 && c[2].vn >= 0;
        if (t.has_vertex_normals)
            for (int k = 0; k < 3; k++) vs[k]->normal = vec(d.vn, c[k].vn);
        t.material_id = tri_material[i];
        scene.AddTriangle(t);
        if (t.material_id >= 0) {
            const Vec3& e = d.materials.list[size_t(t.material_id)].emission;
            if (e.x != 0.0f || e.y != 0.0f || e.z != 0.0f) {
                AreaLight al;
                al.triangle_id = int(scene.Triangles().size()) - 1;
                al.color = e;
                al.surface_area = CalculateTriangleSurfaceArea(t);
                scene.AddAreaLight(al);
            }
        }
    }
    scene.ReserveMaterials(d.materials.list.size());
    for (Material m : d.materials.list) {
        m.r_ior = 1.0f / m.ior;
        scene.AddMaterial(std::move(m));
    }
    return true;
}

bool ObjSceneLoader::LoadMetadata(const std::filesystem::path& metadata_file_path, Scene& scene) {
    std::ifstream in(metadata_file_path);
    if (!in) return Fail("cannot read " + metadata_file_path.string());
    in.imbue(std::locale::classic());
    Camera& cam = scene.GetCamera();
    struct Slot {
        const char* key;
        float* dst[3];
        int n;
    };
    const Slot slots[] = {{"cam_pos", {&cam.pos.x, &cam.pos.y, &cam.pos.z}, 3},
                          {"cam_up", {&cam.up.x, &cam.up.y, &cam.up.z}, 3},
                          {"cam_dir", {&cam.dir.x, &cam.dir.y, &cam.dir.z}, 3},
                          {"cam_fov", {&cam.fov, nullptr, nullptr}, 1},
                          {"cam_far", {&cam.far, nullptr, nullptr}, 1}};
    std::string token;
    while (in >> token) {
        const Slot* s = nullptr;
        for (const Slot& c : slots)
            if (token == c.key) s = &c;
        if (s) {
            for (int i = 0; i < s->n; i++)
                if (!(in >> *s->dst[i])) return Fail(metadata_file_path.string() + ": " + token + " needs numbers");
        } else if (token == "point_light") {
            PointLight pl;
            if (!(in >> pl.pos.x >> pl.pos.y >> pl.pos.z >> pl.color.x >> pl.color.y >> pl.color.z >> pl.power))
                return Fail(metadata_file_path.string() + ": point_light needs 7 numbers");
            scene.AddPointLight(pl);
        } else {
            return Fail(metadata_file_path.string() + ": unrecognized metadata token '" + token + "'");
        }
    }
    return true;
}

std::optional<Scene> ObjSceneLoader::Load(const SceneLoadConfig& config) const {
    const std::filesystem::path base = config.dirpath / config.name;
    Scene scene;
    if (!LoadMetadata(std::filesystem::path(base.string() + ".meta"), scene)) return std::nullopt;
    if (!LoadObj(std::filesystem::path(base.string() + ".obj"), scene)) return std::nullopt;
    return scene;
}

std::unique_ptr<SceneLoader> CreateSceneLoader(const SceneLoadConfig& config) {
    switch (config.scene_load_format) {
        case SceneLoadFormat::kObj:
            return std::make_unique<ObjSceneLoader>();
    }
    throw std::invalid_argument("Unknown SceneLoadFormat");
}

}  // namespace diplodocus_amd

// ---------------------------------------------------------------- C ABI
struct dpl_scene {
    diplodocus_amd::Scene scene;
};

namespace {

// Runs a loader; maps its failure (reason already recorded) and exceptions to a status.
template <class F>
dpl_status Guarded(F&& f) {
    try {
        return f() ? DPL_OK : DPL_ERR_INVALID_ARG;
    } catch (const std::bad_alloc&) {
        diplodocus_amd::t_error = "out of host memory";
        return DPL_ERR_OOM;
    } catch (const std::exception& e) {
        diplodocus_amd::t_error = e.what();
        return DPL_ERR_INTERNAL;
    }
}

}  // namespace

extern "C" {

dpl_status dpl_scene_load(const char* dirpath, const char* name, dpl_scene** out) {
    if (!dirpath || !name || !out) {
        diplodocus_amd::t_error = "dpl_scene_load: NULL argument";
        return DPL_ERR_INVALID_ARG;
    }
    *out = nullptr;
    return Guarded([&] {
        diplodocus_amd::SceneLoadConfig cfg;
        cfg.dirpath = dirpath;
        cfg.name = name;
        auto s = diplodocus_amd::ObjSceneLoader().Load(cfg);
        if (!s) return false;
        *out = new dpl_scene{std::move(*s)};
        return true;
    });
}

dpl_status dpl_scene_load_obj(const char* obj_path, dpl_scene** out) {
    if (!obj_path || !out) {
        diplodocus_amd::t_error = "dpl_scene_load_obj: NULL argument";
        return DPL_ERR_INVALID_ARG;
    }
    *out = nullptr;
    return Guarded([&] {
        auto s = std::make_unique<dpl_scene>();
        if (!diplodocus_amd::ObjSceneLoader::LoadObj(obj_path, s->scene)) return false;
        *out = s.release();
        return true;
    });
}

void dpl_scene_destroy(dpl_scene* scene) { delete scene; }

const char* dpl_scene_last_error(void) { return diplodocus_amd::t_error.c_str(); }

dpl_status dpl_scene_get_info(const dpl_scene* s, dpl_scene_info* info) {
    if (!s || !info) return DPL_ERR_INVALID_ARG;
    const auto& sc = s->scene;
    *info = dpl_scene_info{};
    info->triangle_count = int64_t(sc.Triangles().size());
    info->material_count = int32_t(sc.Materials().size());
    info->area_light_count = int32_t(sc.AreaLights().size());
    info->point_light_count = int32_t(sc.PointLights().size());
    const auto& c = sc.GetCamera();
    const float cam[11] = {c.pos.x, c.pos.y, c.pos.z, c.dir.x, c.dir.y, c.dir.z, c.up.x, c.up.y, c.up.z, c.fov, c.far};
    std::memcpy(info->camera, cam, sizeof(cam));
    return DPL_OK;
}

dpl_status dpl_scene_get_triangles(const dpl_scene* s, float* xyz, float* vertex_normals, uint8_t* has_vertex_normals,
                                   int32_t* material_id, float* geom_normal) {
    if (!s) return DPL_ERR_INVALID_ARG;
    size_t i = 0;
    for (const auto& t : s->scene.Triangles()) {
        const diplodocus_amd::Vertex* v[3] = {&t.v0, &t.v1, &t.v2};
        for (int k = 0; k < 3; k++) {
            if (xyz) xyz[9 * i + 3 * k] = v[k]->pos.x, xyz[9 * i + 3 * k + 1] = v[k]->pos.y, xyz[9 * i + 3 * k + 2] = v[k]->pos.z;
            if (vertex_normals) {
                const diplodocus_amd::Vec3 n = t.has_vertex_normals ? v[k]->normal : diplodocus_amd::Vec3{};
                vertex_normals[9 * i + 3 * k] = n.x, vertex_normals[9 * i + 3 * k + 1] = n.y;
                vertex_normals[9 * i + 3 * k + 2] = n.z;
            }
        }
        if (has_vertex_normals) has_vertex_normals[i] = t.has_vertex_normals ? 1 : 0;
        if (material_id) material_id[i] = t.material_id;
        if (geom_normal) geom_normal[3 * i] = t.geom_normal.x, geom_normal[3 * i + 1] = t.geom_normal.y, geom_normal[3 * i + 2] = t.geom_normal.z;
        i++;
    }
    return DPL_OK;
}

dpl_status dpl_scene_get_materials(const dpl_scene* s, float* rows, dpl_material* render) {
    if (!s) return DPL_ERR_INVALID_ARG;
    size_t i = 0;
    for (const auto& m : s->scene.Materials()) {
        if (rows) {
            const float r[19] = {m.ambient.x,       m.ambient.y,       m.ambient.z,       m.diffuse.x,  m.diffuse.y,
                                 m.diffuse.z,       m.specular.x,      m.specular.y,      m.specular.z, m.transmittance.x,
                                 m.transmittance.y, m.transmittance.z, m.emission.x,      m.emission.y, m.emission.z,
                                 m.shininess,       m.ior,             m.r_ior,           m.dissolve};
            std::memcpy(rows + 19 * i, r, sizeof(r));
        }
        if (render)
            render[i] = dpl_material{{m.diffuse.x, m.diffuse.y, m.diffuse.z},
                                     {m.specular.x, m.specular.y, m.specular.z},
                                     {m.transmittance.x, m.transmittance.y, m.transmittance.z},
                                     {m.emission.x, m.emission.y, m.emission.z},
                                     m.shininess,
                                     m.ior,
                                     m.r_ior,
                                     0.0f};
        i++;
    }
    return DPL_OK;
}

dpl_status dpl_scene_get_material_name(const dpl_scene* s, int32_t material, char* buf, int32_t buf_size) {
    if (!s || !buf || buf_size <= 0 || material < 0 || size_t(material) >= s->scene.Materials().size())
        return DPL_ERR_INVALID_ARG;
    const std::string& n = s->scene.Materials()[size_t(material)].name;
    const size_t len = std::min(n.size(), size_t(buf_size) - 1);
    std::memcpy(buf, n.data(), len);
    buf[len] = '\0';
    return DPL_OK;
}

dpl_status dpl_scene_get_area_lights(const dpl_scene* s, int32_t* triangle_id, float* color, float* surface_area) {
    if (!s) return DPL_ERR_INVALID_ARG;
    size_t i = 0;
    for (const auto& a : s->scene.AreaLights()) {
        if (triangle_id) triangle_id[i] = a.triangle_id;
        if (color) color[3 * i] = a.color.x, color[3 * i + 1] = a.color.y, color[3 * i + 2] = a.color.z;
        if (surface_area) surface_area[i] = a.surface_area;
        i++;
    }
    return DPL_OK;
}

dpl_status dpl_scene_get_point_lights(const dpl_scene* s, float* rows) {
    if (!s) return DPL_ERR_INVALID_ARG;
    size_t i = 0;
    for (const auto& p : s->scene.PointLights()) {
        if (rows) {
            const float r[7] = {p.pos.x, p.pos.y, p.pos.z, p.color.x, p.color.y, p.color.z, p.power};
            std::memcpy(rows + 7 * i, r, sizeof(r));
        }
        i++;
    }
    return DPL_OK;
}

}  // extern "C"
