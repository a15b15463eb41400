"""The reference's JSON configuration from Python (src/io/config/json_config_loader.cc:23-139;
src/config/*.h): the same sections, keys and defaults as the C++ JsonConfigLoader of
diplodocus_amd/host/config.h, with the same rules: a key is read only when present and not
null; numbers convert between int and float (booleans count as 0 / 1), a boolean key needs a
boolean, a Vec3 is an array of at least three numbers, enums are parsed from their names; a
document or section that is not an object keeps every default; malformed JSON, a wrong type
or an unknown enum name raises ConfigError. The acceleration section is an
accel.AccelerationStructureConfig, ready for accel.bridge_acceleration_config."""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field

from .accel import AccelerationStructureConfig, AccelerationStructureType


class ConfigError(RuntimeError):
    pass


def _enum(table: dict, name: str):
    def parse(s, key=None):
        if not isinstance(s, str):
            raise ConfigError(f"expected a string for {name}")
        if s not in table:
            raise ConfigError(f"Invalid {name}: {s}")
        return table[s]

    return parse


parse_renderer_type = _enum({"cpu": "cpu", "gpu": "gpu"}, "RendererType")
parse_image_export_format = _enum({"ppm": "ppm"}, "ImageExportFormat")
parse_stats_export_format = _enum({"console": "console", "csv": "csv"}, "StatsExportFormat")
parse_scene_load_format = _enum({"obj": "obj"}, "SceneLoadFormat")


@dataclass
class ImageExportConfig:  # image_export_config.h:17-20
    image_export_format: str = "ppm"
    filepath: str = "output.ppm"


@dataclass
class RenderConfig:  # render_config.h:18-28
    renderer_type: str = "cpu"
    width: int = 800
    height: int = 600
    background_color: tuple = (0.0, 0.0, 0.0)
    backface_culling: bool = False
    max_depth: int = 4
    area_light_sample_cnt: int = 8
    pixel_sample_cnt: int = 4
    seed: int = 42


@dataclass
class SceneLoadConfig:  # scene_load_config.h:19-24
    scene_load_format: str = "obj"
    dirpath: str = "res"
    name: str = "CornellBox-Sphere"
    triangulate: bool = True


@dataclass
class StatsExportConfig:  # stats_export_config.h:17-19, plus the CSV file
    stats_export_format: str = "console"
    filepath: str = ""


@dataclass
class Config:  # config.h:11-17
    acceleration_structure_config: AccelerationStructureConfig = field(default_factory=AccelerationStructureConfig)
    image_export_config: ImageExportConfig = field(default_factory=ImageExportConfig)
    render_config: RenderConfig = field(default_factory=RenderConfig)
    scene_load_config: SceneLoadConfig = field(default_factory=SceneLoadConfig)
    stats_export_config: StatsExportConfig = field(default_factory=StatsExportConfig)


def _number(v, key, kind):
    if isinstance(v, bool):
        return kind(int(v))
    if isinstance(v, (int, float)):
        if kind is int:
            if isinstance(v, float) and not math.isfinite(v):
                raise ConfigError(f"key '{key}': not a finite number")
            return int(v)  # truncation toward zero, as a C++ conversion
        return float(v)
    raise ConfigError(f"key '{key}': expected a number")


def _as_int(v, key):
    return _number(v, key, int)


def _as_bool(v, key):
    if not isinstance(v, bool):
        raise ConfigError(f"key '{key}': expected a boolean")
    return v


def _as_str(v, key):
    if not isinstance(v, str):
        raise ConfigError(f"key '{key}': expected a string")
    return v


def _as_vec3(v, key):
    if not isinstance(v, list) or len(v) < 3:
        raise ConfigError(f"key '{key}': expected an array of three numbers")
    return tuple(_number(x, key, float) for x in v[:3])


def _section(root, name):
    if not isinstance(root, dict):
        return None
    s = root.get(name)
    return s if isinstance(s, dict) else ({} if s is not None else None)


def _set(sec: dict, obj, key: str, conv):
    """SetIfExists (json_config_loader.cc:23-28): only a present, non-null key is read."""
    if key in sec and sec[key] is not None:
        setattr(obj, key, conv(sec[key], key))


def _reject_constant(name):
    raise ConfigError(f"invalid JSON literal {name}")


def _check_strings(v):
    """Strings must be Unicode text: a lone UTF-16 surrogate escape is an error."""
    if isinstance(v, str):
        if any(0xD800 <= ord(ch) <= 0xDFFF for ch in v):
            raise ConfigError("unpaired surrogate in a string")
    elif isinstance(v, list):
        for x in v:
            _check_strings(x)
    elif isinstance(v, dict):
        for k, x in v.items():
            _check_strings(k)
            _check_strings(x)


def loads(text: str) -> Config:
    """JsonConfigLoader on JSON text."""
    if text.startswith("\ufeff"):
        text = text[1:]
    try:
        root = json.loads(text, parse_constant=_reject_constant)
    except ValueError as e:
        raise ConfigError(f"JSON parse error: {e}") from e
    _check_strings(root)
    c = Config()
    s = _section(root, "image_export")
    if s is not None:
        _set(s, c.image_export_config, "image_export_format", parse_image_export_format)
        _set(s, c.image_export_config, "filepath", _as_str)
    s = _section(root, "render")
    if s is not None:
        r = c.render_config
        _set(s, r, "renderer_type", parse_renderer_type)
        for k in ("width", "height", "max_depth", "area_light_sample_cnt", "pixel_sample_cnt", "seed"):
            _set(s, r, k, _as_int)
        _set(s, r, "background_color", _as_vec3)
        _set(s, r, "backface_culling", _as_bool)
    s = _section(root, "scene_load")
    if s is not None:
        _set(s, c.scene_load_config, "dirpath", _as_str)
        _set(s, c.scene_load_config, "name", _as_str)
        _set(s, c.scene_load_config, "triangulate", _as_bool)
        _set(s, c.scene_load_config, "scene_load_format", parse_scene_load_format)
    s = _section(root, "stats_export")
    if s is not None:
        _set(s, c.stats_export_config, "stats_export_format", parse_stats_export_format)
        _set(s, c.stats_export_config, "filepath", _as_str)
    s = _section(root, "acceleration_structure")
    if s is not None:
        a = c.acceleration_structure_config

        def parse_type(v, key=None):
            try:
                return AccelerationStructureType.parse(_as_str(v, "acceleration_structure_type"))
            except RuntimeError as e:
                raise ConfigError(str(e)) from e

        _set(s, a, "acceleration_structure_type", parse_type)
        for k in ("max_depth", "max_triangles_per_leaf", "nn_search_radius", "kdop_size", "bin_count"):
            _set(s, a, k, _as_int)
    return c


def load(path: str) -> Config:
    """JsonConfigLoader::Load (json_config_loader.cc:43-83)."""
    try:
        with open(path, "r", encoding="utf-8") as f:
            text = f.read()
    except OSError as e:
        raise ConfigError(f"could not open '{path}': {e}") from e
    return loads(text)
