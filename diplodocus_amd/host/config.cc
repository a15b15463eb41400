// JSON configuration and statistics export (config.h).
#include "config.h"

#include <charconv>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <utility>
#include <vector>

namespace diplodocus_amd {

namespace {

thread_local std::string t_config_error;

// ---------------------------------------------------------------- JSON values
struct Json {
    enum class Kind { kNull, kBool, kInt, kFloat, kString, kArray, kObject } kind = Kind::kNull;
    bool b = false;
    int64_t i = 0;
    double f = 0.0;
    std::string s;
    std::vector<Json> items;                           // array elements
    std::vector<std::pair<std::string, Json>> fields;  // object members (a repeated key: last wins)

    const Json* Find(std::string_view key) const {
        if (kind != Kind::kObject) return nullptr;
        for (auto it = fields.rbegin(); it != fields.rend(); ++it)
            if (it->first == key) return &it->second;
        return nullptr;
    }
};

class JsonReader {
   public:
    explicit JsonReader(std::string_view t) : t_(t) {}

    Json Document() {
        if (t_.size() >= 3 && t_.substr(0, 3) == "\xEF\xBB\xBF") p_ = 3;
        Json v = Value(0);
        Space();
        if (p_ != t_.size()) Error("unexpected text after the JSON value");
        return v;
    }

   private:
    [[noreturn]] void Error(const std::string& what) const {
        throw std::runtime_error("JSON parse error at offset " + std::to_string(p_) + ": " + what);
    }
    void Space() {
        while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\t' || t_[p_] == '\n' || t_[p_] == '\r')) p_++;
    }
    bool Eat(char c) {
        Space();
        if (p_ < t_.size() && t_[p_] == c) {
            p_++;
            return true;
        }
        return false;
    }
    void Expect(char c) {
        if (!Eat(c)) Error(std::string("expected '") + c + "'");
    }
    void Literal(std::string_view w) {
        if (t_.substr(p_, w.size()) != w) Error("invalid literal");
        p_ += w.size();
    }

    Json Value(int depth) {
        if (depth > 256) Error("nesting too deep");
        Space();
        if (p_ >= t_.size()) Error("unexpected end of input");
        Json v;
        const char c = t_[p_];
        if (c == '{') {
            p_++;
            v.kind = Json::Kind::kObject;
            if (Eat('}')) return v;
            do {
                Space();
                if (p_ >= t_.size() || t_[p_] != '"') Error("expected a member name");
                std::string key = String();
                Expect(':');
                v.fields.emplace_back(std::move(key), Value(depth + 1));
            } while (Eat(','));
            Expect('}');
        } else if (c == '[') {
            p_++;
            v.kind = Json::Kind::kArray;
            if (Eat(']')) return v;
            do v.items.push_back(Value(depth + 1));
            while (Eat(','));
            Expect(']');
        } else if (c == '"') {
            v.kind = Json::Kind::kString;
            v.s = String();
        } else if (c == 't') {
            Literal("true");
            v.kind = Json::Kind::kBool, v.b = true;
        } else if (c == 'f') {
            Literal("false");
            v.kind = Json::Kind::kBool, v.b = false;
        } else if (c == 'n') {
            Literal("null");
        } else {
            Number(v);
        }
        return v;
    }

    // -?(0|[1-9][0-9]*)(.[0-9]+)?([eE][+-]?[0-9]+)?; integers that fit int64 stay integers
    void Number(Json& v) {
        const size_t s = p_;
        auto digits = [&] {
            const size_t d = p_;
            while (p_ < t_.size() && t_[p_] >= '0' && t_[p_] <= '9') p_++;
            return p_ - d;
        };
        if (p_ < t_.size() && t_[p_] == '-') p_++;
        const size_t int_start = p_;
        const size_t nd = digits();
        if (nd == 0 || (nd > 1 && t_[int_start] == '0')) Error("invalid number");
        bool integral = true;
        if (p_ < t_.size() && t_[p_] == '.') {
            p_++;
            if (digits() == 0) Error("invalid number");
            integral = false;
        }
        if (p_ < t_.size() && (t_[p_] == 'e' || t_[p_] == 'E')) {
            p_++;
            if (p_ < t_.size() && (t_[p_] == '+' || t_[p_] == '-')) p_++;
            if (digits() == 0) Error("invalid number");
            integral = false;
        }
        const char* b = t_.data() + s;
        const char* e = t_.data() + p_;
        if (integral) {
            int64_t x = 0;
            const auto r = std::from_chars(b, e, x);
            if (r.ec == std::errc() && r.ptr == e) {
                v.kind = Json::Kind::kInt, v.i = x;
                return;
            }
        }
        double d = 0.0;
        const auto r = std::from_chars(b, e, d);
        if (r.ptr != e || (r.ec != std::errc() && r.ec != std::errc::result_out_of_range)) Error("invalid number");
        if (r.ec == std::errc::result_out_of_range) Error("number out of range");
        v.kind = Json::Kind::kFloat, v.f = d;
    }

    uint32_t Hex4() {
        if (p_ + 4 > t_.size()) Error("truncated \\u escape");
        uint32_t x = 0;
        for (int k = 0; k < 4; k++) {
            const char c = t_[p_++];
            x <<= 4;
            if (c >= '0' && c <= '9') x |= uint32_t(c - '0');
            else if (c >= 'a' && c <= 'f') x |= uint32_t(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') x |= uint32_t(c - 'A' + 10);
            else Error("invalid \\u escape");
        }
        return x;
    }

    static void Utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) {
            out += char(cp);
        } else if (cp < 0x800) {
            out += char(0xC0 | (cp >> 6)), out += char(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += char(0xE0 | (cp >> 12)), out += char(0x80 | ((cp >> 6) & 0x3F)), out += char(0x80 | (cp & 0x3F));
        } else {
            out += char(0xF0 | (cp >> 18)), out += char(0x80 | ((cp >> 12) & 0x3F));
            out += char(0x80 | ((cp >> 6) & 0x3F)), out += char(0x80 | (cp & 0x3F));
        }
    }

    std::string String() {
        p_++;  // opening quote
        std::string out;
        for (;;) {
            if (p_ >= t_.size()) Error("unterminated string");
            const char c = t_[p_++];
            if (c == '"') return out;
            if (static_cast<unsigned char>(c) < 0x20) Error("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p_ >= t_.size()) Error("unterminated escape");
            const char e = t_[p_++];
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t cp = Hex4();
                    if (cp >= 0xD800 && cp <= 0xDBFF) {
                        if (t_.substr(p_, 2) != "\\u") Error("unpaired surrogate");
                        p_ += 2;
                        const uint32_t lo = Hex4();
                        if (lo < 0xDC00 || lo > 0xDFFF) Error("unpaired surrogate");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (cp >= 0xDC00 && cp <= 0xDFFF) {
                        Error("unpaired surrogate");
                    }
                    Utf8(out, cp);
                    break;
                }
                default:
                    Error("invalid escape");
            }
        }
    }

    std::string_view t_;
    size_t p_ = 0;
};

// ---------------------------------------------------------------- typed reads
[[noreturn]] void TypeError(std::string_view key, std::string_view want) {
    throw std::runtime_error("key '" + std::string(key) + "': expected " + std::string(want));
}

// Numbers (and booleans, as 0 / 1) convert to any arithmetic type.
template <class T>
T Arithmetic(const Json& v, std::string_view key) {
    switch (v.kind) {
        case Json::Kind::kInt: return static_cast<T>(v.i);
        case Json::Kind::kFloat: return static_cast<T>(v.f);
        case Json::Kind::kBool: return static_cast<T>(v.b);
        default: TypeError(key, "a number");
    }
}

// Reads section[key] into out when it is present and not null (SetIfExists).
template <class F>
void IfSet(const Json& section, std::string_view key, F&& read) {
    const Json* v = section.Find(key);
    if (!v || v->kind == Json::Kind::kNull) return;
    read(*v);
}

void SetInt(const Json& sec, std::string_view key, int& out) {
    IfSet(sec, key, [&](const Json& v) { out = Arithmetic<int>(v, key); });
}

void SetBool(const Json& sec, std::string_view key, bool& out) {
    IfSet(sec, key, [&](const Json& v) {
        if (v.kind != Json::Kind::kBool) TypeError(key, "a boolean");
        out = v.b;
    });
}

std::string StringOf(const Json& v, std::string_view key) {
    if (v.kind != Json::Kind::kString) TypeError(key, "a string");
    return v.s;
}

void SetPath(const Json& sec, std::string_view key, std::filesystem::path& out) {
    IfSet(sec, key, [&](const Json& v) { out = StringOf(v, key); });
}

void SetString(const Json& sec, std::string_view key, std::string& out) {
    IfSet(sec, key, [&](const Json& v) { out = StringOf(v, key); });
}

void SetVec3(const Json& sec, std::string_view key, Vec3& out) {
    IfSet(sec, key, [&](const Json& v) {
        if (v.kind != Json::Kind::kArray || v.items.size() < 3) TypeError(key, "an array of three numbers");
        out = {Arithmetic<float>(v.items[0], key), Arithmetic<float>(v.items[1], key), Arithmetic<float>(v.items[2], key)};
    });
}

template <class E, class P>
void SetEnum(const Json& sec, std::string_view key, E& out, P parse) {
    IfSet(sec, key, [&](const Json& v) { out = parse(StringOf(v, key)); });
}

Config FromJson(const Json& root) {
    Config c;
    // json_config_loader.cc:54-75: a section is read when the document is an object holding it
    if (const Json* s = root.Find("image_export")) {
        SetEnum(*s, "image_export_format", c.image_export_config.image_export_format, ParseImageExportFormat);
        SetPath(*s, "filepath", c.image_export_config.filepath);
    }
    if (const Json* s = root.Find("render")) {  // :88-102
        RenderConfig& r = c.render_config;
        SetEnum(*s, "renderer_type", r.renderer_type, ParseRendererType);
        SetInt(*s, "width", r.width);
        SetInt(*s, "height", r.height);
        SetVec3(*s, "background_color", r.background_color);
        SetBool(*s, "backface_culling", r.backface_culling);
        SetInt(*s, "max_depth", r.max_depth);
        SetInt(*s, "area_light_sample_cnt", r.area_light_sample_cnt);
        SetInt(*s, "pixel_sample_cnt", r.pixel_sample_cnt);
        SetInt(*s, "seed", r.seed);
    }
    if (const Json* s = root.Find("scene_load")) {  // :104-113
        SceneLoadConfig& l = c.scene_load_config;
        SetPath(*s, "dirpath", l.dirpath);
        SetString(*s, "name", l.name);
        SetBool(*s, "triangulate", l.triangulate);
        SetEnum(*s, "scene_load_format", l.scene_load_format, ParseSceneLoadFormat);
    }
    if (const Json* s = root.Find("stats_export")) {  // :115-121, plus the CSV file
        SetEnum(*s, "stats_export_format", c.stats_export_config.stats_export_format, ParseStatsExportFormat);
        SetPath(*s, "filepath", c.stats_export_config.filepath);
    }
    if (const Json* s = root.Find("acceleration_structure")) {  // :123-135
        AccelerationStructureConfig& a = c.acceleration_structure_config;
        SetEnum(*s, "acceleration_structure_type", a.acceleration_structure_type, ParseAccelerationStructureType);
        SetInt(*s, "max_depth", a.max_depth);
        SetInt(*s, "max_triangles_per_leaf", a.max_triangles_per_leaf);
        SetInt(*s, "nn_search_radius", a.nn_search_radius);
        SetInt(*s, "kdop_size", a.kdop_size);
        SetInt(*s, "bin_count", a.bin_count);
    }
    return c;
}

// printf "%.3g": the reference's "{:.3}" (general format, three significant digits)
std::string G3(double v) {
    char b[64];
    std::snprintf(b, sizeof(b), "%.3g", v);
    return b;
}

}  // namespace

RendererType ParseRendererType(std::string_view sv) {
    if (sv == "cpu") return RendererType::kCpu;
    if (sv == "gpu") return RendererType::kGpu;
    throw std::runtime_error("Invalid RendererType: " + std::string(sv));
}

ImageExportFormat ParseImageExportFormat(std::string_view sv) {
    if (sv == "ppm") return ImageExportFormat::kPpm;
    throw std::runtime_error("Invalid ImageExportFormat: " + std::string(sv));
}

StatsExportFormat ParseStatsExportFormat(std::string_view sv) {
    if (sv == "console") return StatsExportFormat::kConsole;
    if (sv == "csv") return StatsExportFormat::kCsv;
    throw std::runtime_error("Invalid StatsExportFormat: " + std::string(sv));
}

const std::string& JsonConfigLoader::LastError() { return t_config_error; }

std::optional<Config> JsonConfigLoader::LoadFromString(std::string_view json_text) const {
    try {
        return FromJson(JsonReader(json_text).Document());
    } catch (const std::exception& e) {
        t_config_error = e.what();
        return std::nullopt;
    }
}

std::optional<Config> JsonConfigLoader::Load(const std::filesystem::path& config_file_path) const {
    std::ifstream in(config_file_path, std::ios::binary);
    if (!in) {
        t_config_error = "could not open '" + config_file_path.string() + "'";
        return std::nullopt;
    }
    std::ostringstream ss;
    ss << in.rdbuf();
    auto c = LoadFromString(ss.str());
    if (!c) t_config_error = config_file_path.string() + ": " + t_config_error;
    return c;
}

// ---------------------------------------------------------------- stats export
void ConsoleStatsExporter::ExportOne([[maybe_unused]] const StatsExportConfig& config, const Stats& stats) {
    std::ostream& o = out_ ? *out_ : std::cout;
    const RaytracingStats& r = stats.rt_stats;
    const AccelerationStats& a = stats.accel_stats;
    const float per_query_tests =
        a.query_count == 0 || a.intersection_count == 0 ? 0.0f : float(a.intersection_count) / float(a.query_count);
    const float per_query_nodes =
        a.query_count == 0 || a.traversal_count == 0 ? 0.0f : float(a.traversal_count) / float(a.query_count);
    const char* rule = "########################################\n";
    const char* line = "----------------------------------------\n";
    o << rule << "#                                      #\n"
      << "#              Statistics              #\n"
      << "#                                      #\n"
      << rule << "\n"
      << line << "               Ray tracing              \n"
      << line << " Ray tracing time:              " << G3(r.raytracing_time) << " s\n"
      << " Primary ray count:             " << r.primary_ray_count << "\n"
      << " Secondary ray count:           " << r.secondary_ray_count << "\n"
      << " Shadow ray count:              " << r.shadow_ray_count << "\n"
      << "\n"
      << line << "               Acceleration             \n"
      << line << "# Construction:                         \n"
      << " Build time:                    " << G3(a.build_time) << " ms\n"
      << " Node count:                    " << a.node_count << "\n"
      << " Inner node count:              " << a.inner_node_count << "\n"
      << " Leaf node count:               " << a.leaf_node_count << "\n"
      << " Memory consumption:            " << G3(a.memory_consumption / 1e3f) << " KB\n"
      << "\n"
      << "# Traversal:                             \n"
      << " Query count:                   " << a.query_count << "\n"
      << " Intersection count:            " << a.intersection_count << "\n"
      << " Intersections per query (avg): " << G3(per_query_tests) << "\n"
      << " Traversal count:               " << a.traversal_count << "\n"
      << " Traversals per query:          " << G3(per_query_nodes) << "\n"
      << "\n"
      << rule;
    o.flush();
}

std::string CsvStatsExporter::Header() {
    return "raytracing_time_s,primary_ray_count,secondary_ray_count,shadow_ray_count,build_time_ms,node_count,"
           "inner_node_count,leaf_node_count,memory_consumption_bytes,query_count,intersection_count,traversal_count";
}

std::string CsvStatsExporter::Row(const Stats& s) {
    const RaytracingStats& r = s.rt_stats;
    const AccelerationStats& a = s.accel_stats;
    char t[2][40];
    std::snprintf(t[0], sizeof(t[0]), "%.9g", r.raytracing_time);
    std::snprintf(t[1], sizeof(t[1]), "%.9g", a.build_time);
    std::ostringstream o;
    o << t[0] << ',' << r.primary_ray_count << ',' << r.secondary_ray_count << ',' << r.shadow_ray_count << ','
      << t[1] << ',' << a.node_count << ',' << a.inner_node_count << ',' << a.leaf_node_count << ','
      << a.memory_consumption << ',' << a.query_count << ',' << a.intersection_count << ',' << a.traversal_count;
    return o.str();
}

void CsvStatsExporter::ExportOne(const StatsExportConfig& config, const Stats& stats) {
    if (!config.filepath.empty()) {
        std::error_code ec;
        const bool fresh = !std::filesystem::exists(config.filepath, ec) || std::filesystem::file_size(config.filepath, ec) == 0;
        std::ofstream f(config.filepath, std::ios::app);
        if (!f) throw std::runtime_error("CsvStatsExporter: cannot open " + config.filepath.string());
        if (fresh) f << Header() << '\n';
        f << Row(stats) << '\n';
        return;
    }
    std::ostream& o = out_ ? *out_ : std::cout;
    o << Header() << '\n' << Row(stats) << '\n';
    o.flush();
}

std::unique_ptr<StatsExporter> CreateStatsExporter(const StatsExportConfig& config, std::ostream* out) {
    switch (config.stats_export_format) {
        case StatsExportFormat::kConsole:
            return std::make_unique<ConsoleStatsExporter>(out);
        case StatsExportFormat::kCsv:
            return std::make_unique<CsvStatsExporter>(out);
    }
    throw std::invalid_argument("Unknown StatsExportFormat");
}

}  // namespace diplodocus_amd
