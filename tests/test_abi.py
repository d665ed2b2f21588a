"""CPU checks of the C-ABI boundary: the built library exports every entry point include/crvpt.h declares, the
ctypes table binds exactly those, and calls fail with a message (never a crash) before initialisation."""
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(REPO, "include", "crvpt.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(crvpt_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    assert "crvpt_render_frame" in names and "crvpt_initialize" in names and len(names) > 40


def test_library_exports_every_declared_symbol():
    from cloudrendering_amd import _lib

    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name
    assert set(_lib.SIGNATURES) == set(_declared())


def test_calls_before_initialize_report_errors():
    from cloudrendering_amd import vpt

    with pytest.raises(RuntimeError, match="not initialised"):
        vpt.set_phase_g(0.3)
    with pytest.raises(RuntimeError, match="not initialised"):
        vpt.get_camera_position()


def test_shim_names_follow_reference_module():
    """Every reference binding on the hot path (Module.cpp:85-182) has a same-named function in the shim."""
    from cloudrendering_amd import vpt

    for name in ["initialize", "cleanup", "render_frame", "load_cloud_file", "load_volume_file",
                 "load_environment_map_from_tensor", "set_use_builtin_environment_map",
                 "set_environment_map_intensity", "set_environment_map_intensity_rgb", "set_scattering_albedo",
                 "set_extinction_scale", "set_extinction_base", "set_phase_g", "get_camera_position",
                 "get_camera_view_matrix", "get_camera_fovy", "set_camera_position", "set_camera_target",
                 "overwrite_camera_view_matrix", "set_camera_fovy", "set_vpt_mode", "set_vpt_mode_from_name",
                 "set_denoiser", "set_denoiser_property", "set_use_lights", "clear_lights", "add_light",
                 "remove_light", "set_light_property", "set_use_headlight", "set_seed_offset",
                 "get_volume_voxel_size", "get_render_bounding_box", "set_max_grid_extent"]:
        assert callable(getattr(vpt, name)), name
