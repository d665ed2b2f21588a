"""ctypes binding of libcrvpt.so, the gfx950 volumetric path tracer (C ABI declared in include/crvpt.h).

The library is built in-tree by cloudrendering_amd/build.py. Loading fails loudly when it is missing: there is no
CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CRVPT_LIBRARY", os.path.join(_PKG_DIR, "libcrvpt.so"))

_c_int = ctypes.c_int
_c_i64 = ctypes.c_int64
_c_u32 = ctypes.c_uint32
_c_u64 = ctypes.c_uint64
_c_f = ctypes.c_float
_c_d = ctypes.c_double
_p = ctypes.c_void_p
_s = ctypes.c_char_p
_pd = ctypes.POINTER(ctypes.c_double)
_pi64 = ctypes.POINTER(ctypes.c_int64)
_pu64 = ctypes.POINTER(ctypes.c_uint64)
_pi = ctypes.POINTER(ctypes.c_int)

# Every exported entry point of include/crvpt.h: name -> (restype, argtypes).
SIGNATURES = {
    "crvpt_last_error": (ctypes.c_char_p, []),
    "crvpt_initialize": (_c_int, []),
    "crvpt_cleanup": (_c_int, []),
    "crvpt_render_frame": (_c_int, [_p, _c_int, _c_int, _c_int, _c_int, _c_int, _p]),
    "crvpt_load_cloud_file": (_c_int, [_s]),
    "crvpt_read_xyz_file": (_c_int, [_s, _pi, _pd, _p, _c_u64]),
    "crvpt_read_nvdb_file": (_c_int, [_s, _p, _c_u64, _pu64]),
    "crvpt_set_density_field": (_c_int, [_c_int, _c_int, _c_int, _p]),
    "crvpt_set_nanovdb_buffer": (_c_int, [_p, _c_u64]),
    "crvpt_generate_fbm_cloud": (_c_int, [_c_int, _c_int, _c_int, _c_u32, _c_int, _c_f, _c_f, _c_f, _c_f, _c_f,
                                          _c_int]),
    "crvpt_get_density_field": (_c_int, [_p, _c_u64, _c_int]),
    "crvpt_get_nanovdb_buffer": (_c_int, [_p, _c_u64, _pu64]),
    "crvpt_build_nanovdb_from_dense": (_c_int, [_p, _c_int, _c_int, _c_int, _c_d, _pd, _p, _c_u64, _pu64]),
    "crvpt_build_nanovdb_from_dense_typed": (_c_int, [_p, _c_int, _c_int, _c_int, _c_d, _pd, _c_int, _p, _c_u64,
                                                      _pu64]),
    "crvpt_build_fog_volume_sphere": (_c_int, [_c_f, _p, _c_d, _c_d, _p, _c_u64, _pu64]),
    "crvpt_set_sparse_value_type": (_c_int, [_c_int]),
    "crvpt_set_output_foreground_map": (_c_int, [_c_int]),
    "crvpt_set_use_sparse_grid": (_c_int, [_c_int]),
    "crvpt_set_grid_interpolation_type": (_c_int, [_c_int]),
    "crvpt_set_vpt_mode": (_c_int, [_c_i64]),
    "crvpt_set_vpt_mode_from_name": (_c_int, [_s]),
    "crvpt_set_scattering_albedo": (_c_int, [_pd]),
    "crvpt_set_extinction_scale": (_c_int, [_c_d]),
    "crvpt_set_extinction_base": (_c_int, [_pd]),
    "crvpt_set_phase_g": (_c_int, [_c_d]),
    "crvpt_set_seed_offset": (_c_int, [_c_i64]),
    "crvpt_set_max_bounces": (_c_int, [_c_int]),
    "crvpt_set_super_voxel_size": (_c_int, [_c_int]),
    "crvpt_set_use_feature_maps": (_c_int, [ctypes.c_void_p, _c_int]),
    "crvpt_set_feature_map_type": (_c_int, [_c_int]),
    "crvpt_get_feature_map": (_c_int, [_c_int, ctypes.c_void_p, _c_int, _c_int, _c_int, _c_int, ctypes.c_void_p]),
    "crvpt_set_view_projection_matrix_as_previous": (_c_int, []),
    "crvpt_kpn_per_pixel_kernel": (_c_int, [ctypes.c_void_p] * 3 + [_c_int] * 7 + [ctypes.c_void_p]),
    "crvpt_kpn_inv_iter_kernel": (_c_int, [ctypes.c_void_p] * 6 + [_c_int] * 11 + [ctypes.c_void_p]),
    "crvpt_kpn_inv_iter_kernel_noalloc": (_c_int, [ctypes.c_void_p] * 6 + [_c_int] * 11 + [ctypes.c_void_p]),
    "crvpt_get_camera_position": (_c_int, [_pd]),
    "crvpt_get_camera_view_matrix": (_c_int, [_pd]),
    "crvpt_get_camera_fovy": (_c_int, [_pd]),
    "crvpt_set_camera_position": (_c_int, [_pd]),
    "crvpt_set_camera_target": (_c_int, [_pd]),
    "crvpt_overwrite_camera_view_matrix": (_c_int, [_pd]),
    "crvpt_set_camera_fovy": (_c_int, [_c_d]),
    "crvpt_set_use_builtin_environment_map": (_c_int, [_s]),
    "crvpt_load_environment_map_from_buffer": (_c_int, [_p, _c_int, _c_int, _c_int]),
    "crvpt_set_environment_map_intensity": (_c_int, [_c_d]),
    "crvpt_set_environment_map_intensity_rgb": (_c_int, [_pd]),
    "crvpt_set_use_lights": (_c_int, [_c_int]),
    "crvpt_clear_lights": (_c_int, []),
    "crvpt_add_light": (_c_int, []),
    "crvpt_remove_light": (_c_int, [_c_i64]),
    "crvpt_set_light_property": (_c_int, [_c_i64, _s, _s]),
    "crvpt_get_num_lights": (_c_int, [_pi64]),
    "crvpt_set_max_grid_extent": (_c_int, [_c_d]),
    "crvpt_get_render_bounding_box": (_c_int, [_pd]),
    "crvpt_get_volume_voxel_size": (_c_int, [_pi64]),
    "crvpt_set_denoiser": (_c_int, [_s]),
    "crvpt_set_denoiser_property": (_c_int, [_s, _s]),
    "crvpt_transmittance_batch": (_c_int, [_c_int, _c_int, _p, _p, _p, _c_int, _c_u32, _p, _p, _c_int, _p]),
    "crvpt_set_use_majorant_grid": (_c_int, [_c_int]),
    "crvpt_grid_lookup_batch": (_c_int, [_c_int, _p, _p]),
    "crvpt_grid_lookup_world_batch": (_c_int, [_c_int, _p, _p, _p]),
    "crvpt_get_supervoxel_tables": (_c_int, [_pi, _pi, _p, _p, _c_u64]),
    "crvpt_set_launch_config": (_c_int, [_c_int, _c_int, _c_int]),
    "crvpt_get_last_render_stats": (_c_int, [_pd, _pd, _pu64, _pi]),
    "crvpt_set_collect_stats": (_c_int, [_c_int]),
    "crvpt_get_last_render_counters": (_c_int, [_pu64]),
    "crvpt_kpn_init": (_c_int, [_c_u32]),
    "crvpt_kpn_set_parameters": (_c_int, [_p, _c_u64]),
    "crvpt_kpn_get_parameters": (_c_int, [_p, _c_u64]),
    "crvpt_kpn_compute_gradients": (_c_int, [_p, _p, _p, _c_int, _c_int, _c_int, _p, _p]),
    "crvpt_kpn_get_gradients": (_c_int, [_p, _c_u64]),
    "crvpt_kpn_adam_step": (_c_int, [_c_f, _p]),
    "crvpt_kpn_train_step": (_c_int, [_p, _p, _p, _c_int, _c_int, _c_int, _c_f, _p, _p]),
    "crvpt_kpn_denoise": (_c_int, [_p, _p, _p, _c_int, _c_int, _c_int, _p]),
    "crvpt_kpn_get_logits": (_c_int, [_p, _c_int, _c_int, _c_int, _p]),
    "crvpt_kpn_release": (_c_int, []),
    "crvpt_batch_clear": (_c_int, []),
    "crvpt_batch_add_request": (_c_int, [_p, _c_int, _c_int, _c_int, _c_int, _pi]),
    "crvpt_batch_render": (_c_int, [_c_u32, _p]),
    "crvpt_batch_set_launch_config": (_c_int, [_c_int, _c_int, _c_int, _c_int]),
    "crvpt_batch_get_launch_config": (_c_int, [_pi]),
    "crvpt_batch_autotune": (_c_int, [_c_int, _pi]),
    "crvpt_batch_get_stats": (_c_int, [_pd, _pu64, _pi]),
}

_lib = None


class CrvptError(RuntimeError):
    """Raised when a crvpt entry point returns a non-zero status (message from crvpt_last_error)."""


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads libcrvpt.so once and binds every signature; raises if the library is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise CrvptError(
            f"{path} is missing: build it with `python cloudrendering_amd/build.py` (there is no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def call(name: str, *args) -> None:
    """Calls a crvpt entry point and raises CrvptError with the library's message on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.crvpt_last_error().decode("utf-8", "replace")
        raise CrvptError(msg or f"{name} failed with status {rc}")


def dvec(values, n: int):
    """ctypes double[n] from a sequence (the reference's parseVector3 / parseQuaternion size check)."""
    values = list(values)
    if len(values) != n:
        raise CrvptError(f"Expected vector of size {n}.")
    return (ctypes.c_double * n)(*[float(v) for v in values])
