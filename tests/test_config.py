"""The reference's JSON configuration (src/io/config/json_config_loader.cc:23-139) read by the
Python loader (diplodocus_amd.config), on the reference's own sample_config.json and a file
exercising every key; the C++ JsonConfigLoader runs the same fixtures in
tests/cpp/host_api_test.cc. The acceleration section drives the build parameter bridge."""
import os

import pytest

from diplodocus_amd import accel, config

CFG = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config")


def test_reference_sample_config():
    c = config.load(os.path.join(CFG, "sample_config.json"))
    r = c.render_config
    assert (r.renderer_type, r.width, r.height, r.max_depth, r.area_light_sample_cnt) == ("cpu", 512, 512, 8, 8)
    assert r.background_color == (0.0, 0.0, 0.0) and r.pixel_sample_cnt == 4 and r.seed == 42
    a = c.acceleration_structure_config
    assert a.acceleration_structure_type == accel.AccelerationStructureType.kSbvh and a.nn_search_radius == 10
    assert c.scene_load_config.name == "CornellBox-Bunny" and c.scene_load_config.dirpath == "res"
    assert c.image_export_config.filepath == "output.ppm" and c.stats_export_config.stats_export_format == "console"


def test_every_key():
    c = config.load(os.path.join(CFG, "full_config.json"))
    r = c.render_config
    assert (r.renderer_type, r.width, r.height) == ("gpu", 1024, 768)  # 768.9 truncates
    assert r.background_color == (0.25, 0.5, 0.1) and r.backface_culling and r.max_depth == 6
    assert r.area_light_sample_cnt == 16 and r.pixel_sample_cnt == 4 and r.seed == 7  # null keeps the default
    a = c.acceleration_structure_config
    assert a.acceleration_structure_type == accel.AccelerationStructureType.kPlocEmcSobb
    assert (a.max_depth, a.max_triangles_per_leaf, a.nn_search_radius, a.kdop_size, a.bin_count) == (48, 8, 32, 18, 32)
    assert c.scene_load_config.dirpath == "scenes/obj" and not c.scene_load_config.triangulate
    assert c.stats_export_config.stats_export_format == "csv" and c.stats_export_config.filepath == "stats.csv"
    assert c.image_export_config.filepath == "renders/cornell.ppm"


def test_acceleration_section_drives_the_build_bridge():
    c = config.loads('{"acceleration_structure": {"acceleration_structure_type": "ploc_sobb", '
                     '"nn_search_radius": 24, "max_triangles_per_leaf": 2, "kdop_size": 14}}')
    p = accel.bridge_acceleration_config(c.acceleration_structure_config)
    assert (p.morton_type, p.bv_type, p.nn_search_radius, p.max_triangles_per_leaf, p.kdop_size) == (0, 1, 24, 2, 14)


@pytest.mark.parametrize("text", ["", "{", '{"render": {"width": 01}}', '{"render": {"width": "800"}}',
                                  '{"render": {"renderer_type": "vulkan"}}', '{"render": {}} x',
                                  '{"render": {"backface_culling": 1}}', '{"render": {"background_color": [1, 2]}}',
                                  '{"a": [1, 2,]}', "// c\n{}", '{"s": "\\ud800"}', '{"render": {"width": NaN}}',
                                  '{"acceleration_structure": {"acceleration_structure_type": "bvh8"}}'])
def test_rejects(text):
    with pytest.raises(config.ConfigError):
        config.loads(text)


def test_defaults_for_non_objects():
    assert config.loads("[1, 2, 3]") == config.Config()
    c = config.loads('{"render": 5, "acceleration_structure": {"kdop_size": 14.0}}')
    assert c.render_config == config.RenderConfig() and c.acceleration_structure_config.kdop_size == 14
    with pytest.raises(config.ConfigError):
        config.load("/nonexistent/config.json")
