// Configuration and statistics export for the engine's host interface, in the reference's
// shape so that its JSON files and its stats consumers work unchanged:
//   * Config and its parts: src/config/config.h:11-17, image_export_config.h:10-21,
//     stats_export_config.h:9-20; RenderConfig and AccelerationStructureConfig are the ones of
//     dpl_host.h, SceneLoadConfig the one of scene_loader.h;
//   * JsonConfigLoader::Load (src/io/config/json_config_loader.cc:43-139): the sections
//     "image_export", "render", "scene_load", "stats_export", "acceleration_structure", each key
//     read only when present and not null (SetIfExists, :23-28); numbers convert between
//     integer and floating types, enums are parsed from their names, a Vec3 is an array of
//     at least three numbers; a file that cannot be read or parsed, a value of the wrong type
//     or an unknown enum name gives std::nullopt (reason in JsonConfigLoader::LastError());
//   * StatsExporter / ConsoleStatsExporter / CreateStatsExporter
//     (src/io/stats/stats_exporter.h:9-14, console_stats_exporter.cc:10-60,
//     stats_exporter_factory.cc:10-20), plus the CsvStatsExporter the reference declares but
//     does not implement: one header line and one row per export.
// The JSON reader is this library's own (RFC 8259: no comments, no trailing commas; a UTF-8
// byte order mark is skipped; a repeated key keeps its last value).
#pragma once

#include <filesystem>
#include <iosfwd>
#include <memory>
#include <optional>
#include <string>
#include <string_view>

#include "dpl_host.h"
#include "scene_loader.h"

namespace diplodocus_amd {

enum class ImageExportFormat { kPpm };
enum class StatsExportFormat { kConsole, kCsv };

// "cpu" / "gpu", "ppm", "console" / "csv"; throw std::runtime_error on anything else.
RendererType ParseRendererType(std::string_view sv);
ImageExportFormat ParseImageExportFormat(std::string_view sv);
StatsExportFormat ParseStatsExportFormat(std::string_view sv);

struct ImageExportConfig {
    ImageExportFormat image_export_format = ImageExportFormat::kPpm;
    std::filesystem::path filepath = "output.ppm";
};

struct StatsExportConfig {
    StatsExportFormat stats_export_format = StatsExportFormat::kConsole;
    // CSV only: the file rows are appended to (its header is written when the file is new or
    // empty); empty = the exporter's stream.
    std::filesystem::path filepath;
};

struct Config {
    AccelerationStructureConfig acceleration_structure_config;
    ImageExportConfig image_export_config;
    RenderConfig render_config;
    SceneLoadConfig scene_load_config;
    StatsExportConfig stats_export_config;
};

class JsonConfigLoader {
   public:
    std::optional<Config> Load(const std::filesystem::path& config_file_path) const;
    // The same from JSON text (no file).
    std::optional<Config> LoadFromString(std::string_view json_text) const;
    static const std::string& LastError();
};

class StatsExporter {
   public:
    virtual ~StatsExporter() = default;
    virtual void ExportOne(const StatsExportConfig& stats_export_config, const Stats& stats) = 0;
};

// The reference's console report. Its "Traversals per query" line divides the traversal
// count (the reference divides the intersection count there, console_stats_exporter.cc:42-43).
class ConsoleStatsExporter final : public StatsExporter {
   public:
    explicit ConsoleStatsExporter(std::ostream* out = nullptr) : out_(out) {}  // null: std::cout
    void ExportOne(const StatsExportConfig& stats_export_config, const Stats& stats) override;

   private:
    std::ostream* out_;
};

// Columns: raytracing_time_s, primary_ray_count, secondary_ray_count, shadow_ray_count,
// build_time_ms, node_count, inner_node_count, leaf_node_count, memory_consumption_bytes,
// query_count, intersection_count, traversal_count.
class CsvStatsExporter final : public StatsExporter {
   public:
    explicit CsvStatsExporter(std::ostream* out = nullptr) : out_(out) {}  // null: std::cout
    void ExportOne(const StatsExportConfig& stats_export_config, const Stats& stats) override;
    static std::string Header();
    static std::string Row(const Stats& stats);

   private:
    std::ostream* out_;
};

// The factory (stats_exporter_factory.cc:10-20); out as for the exporters.
std::unique_ptr<StatsExporter> CreateStatsExporter(const StatsExportConfig& config, std::ostream* out = nullptr);

}  // namespace diplodocus_amd
