"""Python surface of the gfx950 volumetric path tracer, named and shaped like the reference's `vpt` module.

Every function here follows the reference binding of the same name (src/PyTorch/Module.cpp:85-182, stubs in
src/PyTorch/vpt.pyi) and forwards to the C ABI of libcrvpt.so (include/crvpt.h). PyTorch tensors are used only
to hand memory to the library: a CPU input tensor gives a CPU result (renderFrameCpu), a HIP tensor (device type
"cuda" under ROCm) is rendered in place on the current HIP stream without host round trips (the reference's CUDA
branch). Errors the reference raises through sgl::Logfile::throwError are raised as RuntimeError (CrvptError).
"""
from __future__ import annotations

import ctypes
from typing import List, Sequence

import numpy as np

from . import _lib
from ._lib import CrvptError, call, dvec

__all__ = [
    "initialize", "cleanup", "render_frame", "load_cloud_file", "load_volume_file",
    "load_environment_map_from_tensor", "set_use_builtin_environment_map", "set_environment_map_intensity",
    "set_environment_map_intensity_rgb", "set_scattering_albedo", "set_extinction_scale", "set_extinction_base",
    "set_phase_g", "get_camera_position", "get_camera_view_matrix", "get_camera_fovy", "set_camera_position",
    "set_camera_target", "overwrite_camera_view_matrix", "set_camera_fovy", "set_vpt_mode",
    "set_vpt_mode_from_name", "set_denoiser", "set_denoiser_property", "set_use_lights", "clear_lights",
    "add_light", "remove_light", "set_light_property", "set_use_headlight", "set_use_headlight_distance",
    "set_headlight_color", "set_headlight_intensity", "set_seed_offset", "get_volume_voxel_size",
    "get_render_bounding_box", "set_max_grid_extent",
    # pass settings API (VolumetricPathTracingPass.hpp:136-225) and build-specific extensions
    "set_use_sparse_grid", "set_grid_interpolation_type", "set_density_field", "set_nanovdb_buffer",
    "get_density_field", "get_nanovdb_buffer", "generate_fbm_cloud", "transmittance_batch",
    "set_use_majorant_grid", "set_launch_config", "get_last_render_stats", "get_num_lights",
    "set_collect_stats", "get_last_render_counters", "build_nanovdb_from_dense", "grid_lookup",
    "grid_lookup_world", "get_supervoxel_tables", "set_max_bounces", "set_super_voxel_size",
    "set_use_feature_maps", "get_feature_map_from_string", "get_feature_map", "set_feature_map_type",
    "set_view_projection_matrix_as_previous", "VPT_FEATURE_MAP_NAMES", "per_pixel_kernel_forward", "inv_iter_kernel",
    "VPT_MODE_NAMES", "CrvptError", "batch_clear", "batch_add_request", "batch_render", "batch_set_launch_config",
    "batch_get_launch_config", "batch_autotune", "batch_get_stats", "set_sparse_value_type",
]

# VPT_MODE_NAMES (RenderSettings.hpp:37-41)
VPT_MODE_NAMES = [
    "Delta Tracking", "Delta Tracking (Spectral)", "Ratio Tracking", "Decomposition Tracking",
    "Residual Ratio Tracking", "Next Event Tracking", "Next Event Tracking (Spectral)", "Isosurfaces",
    "Ray Marching (Emission/Absorption)",
]


def initialize() -> None:
    call("crvpt_initialize")


def cleanup() -> None:
    call("crvpt_cleanup")


def _torch():
    import torch  # noqa: PLC0415 - torch is only needed for the tensor-facing entry points

    return torch


def render_frame(input_tensor, frame_count: int):
    """Module.cpp:286-317: renders frame_count samples per pixel and returns a new [C, H, W] float32 tensor on the
    input tensor's device. Only the shape, device and dtype of input_tensor are used."""
    torch = _torch()
    if input_tensor.dim() != 3:
        raise CrvptError("Error in renderFrame: inputTensor.sizes().size() != 3.")
    c, h, w = (int(s) for s in input_tensor.shape)
    if c != 3 and c != 4:
        raise CrvptError("Error in renderFrame: The number of image channels is not equal to 3 or 4.")
    if input_tensor.dtype != torch.float32:
        raise CrvptError("Error in renderFrame: The only data type currently supported is 32-bit float.")
    dev = input_tensor.device
    out = torch.empty((c, h, w), dtype=torch.float32, device=dev)
    if dev.type == "cpu":
        call("crvpt_render_frame", ctypes.c_void_p(out.data_ptr()), c, h, w, int(frame_count), 0, None)
    elif dev.type == "cuda":
        stream = torch.cuda.current_stream(dev).cuda_stream
        call("crvpt_render_frame", ctypes.c_void_p(out.data_ptr()), c, h, w, int(frame_count), 1,
             ctypes.c_void_p(stream))
    else:
        raise CrvptError("Unsupported PyTorch device type.")
    return out


# VPT_FEATURE_MAP_NAMES (VolumetricPathTracingPass.hpp:74-80), indexed by FeatureMapTypeVpt
VPT_FEATURE_MAP_NAMES = [
    "Result", "First X", "First W", "Normal", "Normal (Length 1)", "Cloud Only", "Depth", "Depth Far",
    "Flow", "Flow Reverse", "Depth (nabla)", "Depth (fwidth)", "Density", "Background", "Reprojected UV",
    "Depth Blended", "Depth Nearest Opaque", "Albedo",
    "Transmittance Volume",
    "Primary Ray Absorption Moments", "Scatter Ray Absorption Moments",
    "Radiance Variance",  # extension: per-pixel sample variance of the accumulated radiance
]


def set_use_feature_maps(feature_map_names: Sequence[str]) -> None:
    """Module.cpp:319-330: names not in VPT_FEATURE_MAP_NAMES are ignored, as in the reference."""
    idx = [VPT_FEATURE_MAP_NAMES.index(n) for n in feature_map_names if n in VPT_FEATURE_MAP_NAMES]
    arr = (ctypes.c_int32 * max(1, len(idx)))(*idx)
    call("crvpt_set_use_feature_maps", ctypes.cast(arr, ctypes.c_void_p), len(idx))


def set_feature_map_type(type: int) -> None:  # noqa: A002 - the reference's argument name
    call("crvpt_set_feature_map_type", int(type))


def get_feature_map(input_tensor, feature_map: int):
    """Module.cpp:343-400: the feature map of the last render as a new [C, H, W] float32 tensor on the input
    tensor's device (shape, device and dtype of input_tensor are used, not its contents)."""
    torch = _torch()
    if input_tensor.dim() != 3:
        raise CrvptError("Error in renderFrame: inputTensor.sizes().size() != 3.")
    c, h, w = (int(s) for s in input_tensor.shape)
    if c != 3 and c != 4:
        raise CrvptError("Error in renderFrame: The number of image channels is not equal to 3 or 4.")
    if input_tensor.dtype != torch.float32:
        raise CrvptError("Error in renderFrame: The only data type currently supported is 32-bit float.")
    dev = input_tensor.device
    out = torch.empty((c, h, w), dtype=torch.float32, device=dev)
    if dev.type == "cpu":
        call("crvpt_get_feature_map", int(feature_map), ctypes.c_void_p(out.data_ptr()), c, h, w, 0, None)
    elif dev.type == "cuda":
        stream = torch.cuda.current_stream(dev).cuda_stream
        call("crvpt_get_feature_map", int(feature_map), ctypes.c_void_p(out.data_ptr()), c, h, w, 1,
             ctypes.c_void_p(stream))
    else:
        raise CrvptError("Unsupported PyTorch device type.")
    return out


def get_feature_map_from_string(input_tensor, feature_map: str):
    """Module.cpp:332-341: unknown names select map 0 (Result), as in the reference."""
    idx = VPT_FEATURE_MAP_NAMES.index(feature_map) if feature_map in VPT_FEATURE_MAP_NAMES else 0
    return get_feature_map(input_tensor, idx)


def set_view_projection_matrix_as_previous() -> None:
    call("crvpt_set_view_projection_matrix_as_previous")


def _dev4(t, name):
    torch = _torch()
    if t.dim() != 4:
        raise CrvptError(f"Error in perPixelKernelCuda: {name}.sizes().size() != 4.")
    if t.device.type != "cuda":
        raise CrvptError("Error in perPixelKernelForward: Unsupported device.")
    return t.to(torch.float32).contiguous()


def per_pixel_kernel_forward(image, weights, kernel_size: int):
    """modules/kpn_module per_pixel_kernel::forward: image [N, C<=4, H, W], weights [N, K*K, Hw, Ww] on the GPU ->
    [N, C, min(H, Hw), min(W, Ww)]."""
    torch = _torch()
    image = _dev4(image, "image")
    weights = _dev4(weights, "weights")
    n, c, h, w = (int(v) for v in image.shape)
    wh, ww = int(weights.shape[2]), int(weights.shape[3])
    out = torch.zeros((n, c, min(h, wh), min(w, ww)), dtype=torch.float32, device=image.device)
    stream = torch.cuda.current_stream(image.device).cuda_stream
    call("crvpt_kpn_per_pixel_kernel", ctypes.c_void_p(image.data_ptr()), ctypes.c_void_p(weights.data_ptr()),
         ctypes.c_void_p(out.data_ptr()), n, c, h, w, wh, ww, int(kernel_size), ctypes.c_void_p(stream))
    return out


def inv_iter_kernel(x, low_res_pred, low_res_x, prev, fused_weights, kernel_size: int, use_previous: bool):
    """PyTorchDenoiser.cpp:174-179: blend base (getInvIterBase) then the per-pixel kernel with softmax over the
    first K*K channels of fused_weights."""
    torch = _torch()
    x = _dev4(x, "image")
    low_res_pred = _dev4(low_res_pred, "image")
    low_res_x = _dev4(low_res_x, "image")
    fused_weights = _dev4(fused_weights, "weights")
    prev = _dev4(prev, "image") if prev is not None else x
    n, c, h, w = (int(v) for v in x.shape)
    lh, lw = int(low_res_x.shape[2]), int(low_res_x.shape[3])
    wc, wh, ww = (int(v) for v in fused_weights.shape[1:])
    out = torch.zeros((n, c, min(h, wh), min(w, ww)), dtype=torch.float32, device=x.device)
    stream = torch.cuda.current_stream(x.device).cuda_stream
    call("crvpt_kpn_inv_iter_kernel", ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(low_res_pred.data_ptr()),
         ctypes.c_void_p(low_res_x.data_ptr()), ctypes.c_void_p(prev.data_ptr()),
         ctypes.c_void_p(fused_weights.data_ptr()), ctypes.c_void_p(out.data_ptr()), n, c, h, w, lh, lw, wc, wh, ww,
         int(kernel_size), int(bool(use_previous)), ctypes.c_void_p(stream))
    return out


def load_cloud_file(filename: str) -> None:
    call("crvpt_load_cloud_file", filename.encode())


load_volume_file = load_cloud_file


def load_environment_map_from_tensor(env_map_tensor) -> None:
    """Module.cpp:897-961: [C=3|4, H, W] float tensor on the CPU, linear RGB."""
    torch = _torch()
    if env_map_tensor.device.type != "cpu":
        raise CrvptError("Error in loadEnvironmentMapFromTensor: Expected tensor using CPU memory.")
    t = env_map_tensor.detach().to(torch.float32).contiguous()
    c, h, w = (int(s) for s in t.shape)
    call("crvpt_load_environment_map_from_buffer", ctypes.c_void_p(t.data_ptr()), c, h, w)


def set_use_builtin_environment_map(env_map_name: str) -> None:
    call("crvpt_set_use_builtin_environment_map", env_map_name.encode())


def set_environment_map_intensity(intensity_factor: float) -> None:
    call("crvpt_set_environment_map_intensity", float(intensity_factor))


def set_environment_map_intensity_rgb(intensity_factor: Sequence[float]) -> None:
    call("crvpt_set_environment_map_intensity_rgb", dvec(intensity_factor, 3))


def set_scattering_albedo(albedo: Sequence[float]) -> None:
    call("crvpt_set_scattering_albedo", dvec(albedo, 3))


def set_extinction_scale(extinction_scale: float) -> None:
    call("crvpt_set_extinction_scale", float(extinction_scale))


def set_extinction_base(extinction_base: Sequence[float]) -> None:
    call("crvpt_set_extinction_base", dvec(extinction_base, 3))


def set_phase_g(phase_g: float) -> None:
    call("crvpt_set_phase_g", float(phase_g))


def get_camera_position() -> List[float]:
    v = (ctypes.c_double * 3)()
    call("crvpt_get_camera_position", v)
    return list(v)


def get_camera_view_matrix() -> List[float]:
    v = (ctypes.c_double * 16)()
    call("crvpt_get_camera_view_matrix", v)
    return list(v)


def get_camera_fovy() -> float:
    v = ctypes.c_double()
    call("crvpt_get_camera_fovy", ctypes.byref(v))
    return v.value


def set_camera_position(camera_position: Sequence[float]) -> None:
    call("crvpt_set_camera_position", dvec(camera_position, 3))


def set_camera_target(camera_target: Sequence[float]) -> None:
    call("crvpt_set_camera_target", dvec(camera_target, 3))


def overwrite_camera_view_matrix(view_matrix_data: Sequence[float]) -> None:
    call("crvpt_overwrite_camera_view_matrix", dvec(view_matrix_data, 16))


def set_camera_fovy(fovy: float) -> None:
    """Module.cpp:1384-1387: fovy in degrees."""
    call("crvpt_set_camera_fovy", float(fovy))


def set_vpt_mode(mode: int) -> None:
    call("crvpt_set_vpt_mode", int(mode))


def set_vpt_mode_from_name(mode_name: str) -> None:
    call("crvpt_set_vpt_mode_from_name", mode_name.encode())


def set_denoiser(denoiser_name: str) -> None:
    call("crvpt_set_denoiser", denoiser_name.encode())


def set_denoiser_property(key: str, value: str) -> None:
    call("crvpt_set_denoiser_property", key.encode(), str(value).encode())


def set_output_foreground_map(_shall_output_foreground_map: bool) -> None:
    call("crvpt_set_output_foreground_map", int(bool(_shall_output_foreground_map)))


def set_use_lights(_use_lights: bool) -> None:
    call("crvpt_set_use_lights", int(bool(_use_lights)))


def clear_lights() -> None:
    call("crvpt_clear_lights")


def add_light() -> None:
    call("crvpt_add_light")


def remove_light(light_idx: int) -> None:
    call("crvpt_remove_light", int(light_idx))


def set_light_property(light_idx: int, key: str, value: str) -> None:
    call("crvpt_set_light_property", int(light_idx), key.encode(), str(value).encode())


def get_num_lights() -> int:
    v = ctypes.c_int64()
    call("crvpt_get_num_lights", ctypes.byref(v))
    return int(v.value)


def set_use_headlight(_use_headlight: bool) -> None:
    """Module.cpp:1121-1131: the old headlight API maps onto light 0 in view space."""
    set_use_lights(_use_headlight)
    if _use_headlight:
        clear_lights()
        add_light()
        set_light_property(0, "space", "View")
        set_light_property(0, "position", "0 0 0")
        set_light_property(0, "use_distance", "1")


def set_use_headlight_distance(_use_headlight_distance: bool) -> None:
    set_light_property(0, "use_distance", "1" if _use_headlight_distance else "0")


def set_headlight_color(_headlight_color: Sequence[float]) -> None:
    c = list(_headlight_color)
    if len(c) != 3:
        raise CrvptError("Expected vector of size 3.")
    set_light_property(0, "color", f"{c[0]:f} {c[1]:f} {c[2]:f}")


def set_headlight_intensity(_headlight_intensity: float) -> None:
    set_light_property(0, "intensity", f"{float(_headlight_intensity):f}")


def set_seed_offset(offset: int) -> None:
    call("crvpt_set_seed_offset", int(offset))


def set_super_voxel_size(super_voxel_size: int) -> None:
    """Pass setting superVoxelSize (VolumetricPathTracingPass.hpp:254): edge of the residual-ratio and decomposition
    supervoxel tables, 1-64, default 8."""
    call("crvpt_set_super_voxel_size", int(super_voxel_size))


def set_max_bounces(max_bounces: int) -> None:
    """Scattering events per path before it ends with zero further radiance; 0 (the default) is unbounded, as in
    the reference trackers."""
    call("crvpt_set_max_bounces", int(max_bounces))


def get_volume_voxel_size() -> List[int]:
    v = (ctypes.c_int64 * 3)()
    call("crvpt_get_volume_voxel_size", v)
    return list(v)


def get_render_bounding_box() -> List[float]:
    v = (ctypes.c_double * 6)()
    call("crvpt_get_render_bounding_box", v)
    return list(v)


def set_max_grid_extent(max_grid_extent: float) -> None:
    """The grid spans (-maxGridExtent, maxGridExtent) in its largest dimension; call before loading a volume."""
    call("crvpt_set_max_grid_extent", float(max_grid_extent))


# ---------------------------------------------------------------------------------------------------------------
# Pass settings API and build-specific extensions
# ---------------------------------------------------------------------------------------------------------------

def set_use_sparse_grid(use_sparse: bool) -> None:
    """VolumetricPathTracingPass::setUseSparseGrid: render the NanoVDB representation (built from the dense field
    when none was loaded)."""
    call("crvpt_set_use_sparse_grid", int(bool(use_sparse)))


GRID_INTERPOLATION_TYPE_NAMES = ["Nearest", "Stochastic", "Trilinear"]  # RenderSettings.hpp:57-59


def set_grid_interpolation_type(interpolation) -> None:
    """VolumetricPathTracingPass::setSparseGridInterpolationType: 0 nearest, 1 stochastic, 2 trilinear (or the
    name, case-insensitive)."""
    if isinstance(interpolation, str):
        names = [n.lower() for n in GRID_INTERPOLATION_TYPE_NAMES]
        if interpolation.lower() not in names:
            raise CrvptError(f"set_grid_interpolation_type: unknown type \"{interpolation}\"")
        interpolation = names.index(interpolation.lower())
    call("crvpt_set_grid_interpolation_type", int(interpolation))


def set_density_field(density) -> None:
    """CloudData::setDensityField for a [Z, Y, X] array (x fastest in memory)."""
    a = np.ascontiguousarray(np.asarray(density, dtype=np.float32))
    if a.ndim != 3:
        raise CrvptError("set_density_field: expected a 3D array [Z, Y, X]")
    nz, ny, nx = a.shape
    call("crvpt_set_density_field", nx, ny, nz, a.ctypes.data_as(ctypes.c_void_p))


def set_nanovdb_buffer(buffer: bytes) -> None:
    b = np.frombuffer(bytes(buffer), dtype=np.uint8)
    call("crvpt_set_nanovdb_buffer", b.ctypes.data_as(ctypes.c_void_p), int(b.size))


def get_density_field(shape_zyx: Sequence[int], normalized: bool = False) -> np.ndarray:
    nz, ny, nx = (int(s) for s in shape_zyx)
    out = np.empty((nz, ny, nx), dtype=np.float32)
    call("crvpt_get_density_field", out.ctypes.data_as(ctypes.c_void_p), out.size, int(bool(normalized)))
    return out


def get_nanovdb_buffer() -> bytes:
    size = ctypes.c_uint64()
    call("crvpt_get_nanovdb_buffer", None, 0, ctypes.byref(size))
    out = np.empty(size.value, dtype=np.uint8)
    call("crvpt_get_nanovdb_buffer", out.ctypes.data_as(ctypes.c_void_p), out.size, ctypes.byref(size))
    return out.tobytes()


def generate_fbm_cloud(nx: int, ny: int, nz: int, seed: int = 0, octaves: int = 5, lacunarity: float = 2.0,
                       gain: float = 0.5, base_frequency: float = 4.0, edge0: float = 0.45, edge1: float = 0.6,
                       height_falloff: bool = False) -> None:
    """Generates the seeded fBm procedural cloud on the GPU and makes it the current dense density field."""
    call("crvpt_generate_fbm_cloud", nx, ny, nz, seed, octaves, lacunarity, gain, base_frequency, edge0, edge1,
         int(bool(height_falloff)))


ESTIMATORS = {"delta": 0, "ratio": 1, "residual_ratio": 2, "decomposition": 3}


def transmittance_batch(estimator, origins, directions, t_max=None, n_samples: int = 1, seed: int = 0):
    """Mean and second moment of n_samples transmittance estimates per ray (see include/crvpt.h). Torch HIP
    tensors stay on the device; NumPy arrays / CPU tensors are copied."""
    est = ESTIMATORS[estimator] if isinstance(estimator, str) else int(estimator)
    try:
        torch = _torch()
        is_dev = isinstance(origins, torch.Tensor) and origins.device.type == "cuda"
    except ImportError:  # pragma: no cover - torch is part of the environment
        torch, is_dev = None, False
    if is_dev:
        o = origins.to(torch.float32).contiguous()
        d = directions.to(torch.float32).contiguous()
        tm = t_max.to(torch.float32).contiguous() if t_max is not None else None
        n = int(o.shape[0])
        mean = torch.empty(n, dtype=torch.float32, device=o.device)
        m2 = torch.empty(n, dtype=torch.float32, device=o.device)
        stream = torch.cuda.current_stream(o.device).cuda_stream
        call("crvpt_transmittance_batch", est, n, ctypes.c_void_p(o.data_ptr()), ctypes.c_void_p(d.data_ptr()),
             ctypes.c_void_p(tm.data_ptr()) if tm is not None else None, int(n_samples), int(seed) & 0xFFFFFFFF,
             ctypes.c_void_p(mean.data_ptr()), ctypes.c_void_p(m2.data_ptr()), 1, ctypes.c_void_p(stream))
        return mean, m2
    o = np.ascontiguousarray(np.asarray(origins, dtype=np.float32).reshape(-1, 3))
    d = np.ascontiguousarray(np.asarray(directions, dtype=np.float32).reshape(-1, 3))
    n = o.shape[0]
    tm = np.ascontiguousarray(np.asarray(t_max, dtype=np.float32).reshape(n)) if t_max is not None else None
    mean = np.empty(n, dtype=np.float32)
    m2 = np.empty(n, dtype=np.float32)
    call("crvpt_transmittance_batch", est, n, o.ctypes.data_as(ctypes.c_void_p), d.ctypes.data_as(ctypes.c_void_p),
         tm.ctypes.data_as(ctypes.c_void_p) if tm is not None else None, int(n_samples), int(seed) & 0xFFFFFFFF,
         mean.ctypes.data_as(ctypes.c_void_p), m2.ctypes.data_as(ctypes.c_void_p), 0, None)
    return mean, m2


def set_use_majorant_grid(use_majorant_grid: bool) -> None:
    call("crvpt_set_use_majorant_grid", int(bool(use_majorant_grid)))


def set_launch_config(supervoxel_size: int = 0, blocks: int = 0, steps_per_refill: int = 0) -> None:
    call("crvpt_set_launch_config", int(supervoxel_size), int(blocks), int(steps_per_refill))


# Batched render queue (no reference counterpart). Requests snapshot the current scene as render_frame would render
# it; batch_render traces every request in one persistent launch (a replayed HIP graph).
_batch_outputs: list = []


def batch_clear() -> None:
    call("crvpt_batch_clear")
    _batch_outputs.clear()


def batch_add_request(input_tensor, frame_count: int):
    """Queues frame_count fresh samples per pixel of the current scene; returns (request id, output tensor). The
    output is a new [C, H, W] float32 tensor on the input tensor's GPU (only its shape, device and dtype are used),
    written by batch_render."""
    torch = _torch()
    if input_tensor.dim() != 3 or int(input_tensor.shape[0]) not in (3, 4):
        raise CrvptError("batch_add_request: expected a [3 or 4, H, W] tensor")
    if input_tensor.device.type != "cuda" or input_tensor.dtype != torch.float32:
        raise CrvptError("batch_add_request: expected a float32 tensor on the GPU")
    c, h, w = (int(v) for v in input_tensor.shape)
    out = torch.zeros((c, h, w), dtype=torch.float32, device=input_tensor.device)
    rid = ctypes.c_int()
    call("crvpt_batch_add_request", ctypes.c_void_p(out.data_ptr()), c, h, w, int(frame_count), ctypes.byref(rid))
    _batch_outputs.append(out)
    return rid.value, out


def batch_render(frame_offset: int = 0) -> list:
    """Renders every queued request on torch's current stream; returns the output tensors in request order."""
    torch = _torch()
    if not _batch_outputs:
        raise CrvptError("batch_render: the batch holds no request")
    stream = torch.cuda.current_stream(_batch_outputs[0].device).cuda_stream
    call("crvpt_batch_render", int(frame_offset), ctypes.c_void_p(stream))
    return list(_batch_outputs)


def _batch_config(cfg) -> dict:
    return {"blocks": cfg[0], "block_size": cfg[1], "steps_per_refill": cfg[2], "samples_per_unit": cfg[3]}


def batch_set_launch_config(blocks: int = 0, block_size: int = 256, steps_per_refill: int = 16,
                            samples_per_unit: int = 16) -> None:
    call("crvpt_batch_set_launch_config", int(blocks), int(block_size), int(steps_per_refill), int(samples_per_unit))


def batch_get_launch_config() -> dict:
    cfg = (ctypes.c_int * 4)()
    call("crvpt_batch_get_launch_config", cfg)
    return _batch_config(cfg)


def batch_autotune(spp_divisor: int = 16) -> dict:
    cfg = (ctypes.c_int * 4)()
    call("crvpt_batch_autotune", int(spp_divisor), cfg)
    return _batch_config(cfg)


def batch_get_stats() -> dict:
    ms = ctypes.c_double()
    n = ctypes.c_uint64()
    nreq = ctypes.c_int()
    call("crvpt_batch_get_stats", ctypes.byref(ms), ctypes.byref(n), ctypes.byref(nreq))
    return {"kernel_ms": ms.value, "num_paths": int(n.value), "num_requests": nreq.value}


def get_last_render_stats() -> dict:
    a, b = ctypes.c_double(), ctypes.c_double()
    n = ctypes.c_uint64()
    launches = ctypes.c_int()
    call("crvpt_get_last_render_stats", ctypes.byref(a), ctypes.byref(b), ctypes.byref(n), ctypes.byref(launches))
    return {"path_kernel_ms": a.value, "total_ms": b.value, "num_paths": int(n.value), "launches": launches.value}


def set_collect_stats(collect: bool) -> None:
    call("crvpt_set_collect_stats", int(bool(collect)))


def get_last_render_counters() -> dict:
    v = (ctypes.c_uint64 * 4)()
    call("crvpt_get_last_render_counters", v)
    return {"density_taps": int(v[0]), "supervoxel_reads": int(v[1]), "paths": int(v[2]),
            "tree_descents": int(v[3])}


def build_nanovdb_from_dense(density_zyx, dx: float, origin=(0.0, 0.0, 0.0), value_type: int = 0) -> bytes:
    """Host-side dense -> NanoVDB conversion (CloudData::getSparseDensityField semantics); no GPU needed.
    value_type 0: FP32 leaves (the reference's); 1: Fp16 leaves (16-bit codes per leaf)."""
    a = np.ascontiguousarray(np.asarray(density_zyx, dtype=np.float32))
    nz, ny, nx = a.shape
    o = dvec(origin, 3)
    size = ctypes.c_uint64()
    call("crvpt_build_nanovdb_from_dense_typed", a.ctypes.data_as(ctypes.c_void_p), nx, ny, nz, float(dx), o,
         int(value_type), None, 0, ctypes.byref(size))
    out = np.empty(size.value, dtype=np.uint8)
    call("crvpt_build_nanovdb_from_dense_typed", a.ctypes.data_as(ctypes.c_void_p), nx, ny, nz, float(dx), o,
         int(value_type), out.ctypes.data_as(ctypes.c_void_p), out.size, ctypes.byref(size))
    return out.tobytes()


def build_fog_volume_sphere(radius: float = 0.25, center=(0.0, 0.0, 0.0), voxel_size: float = 0.01,
                            half_width: float = 3.0) -> bytes:
    """nanovdb::createFogVolumeSphere<float> (the default scene's grid, Module.cpp:793-795); no GPU needed."""
    c = (ctypes.c_float * 3)(*[float(v) for v in center])
    vs = float(np.float32(voxel_size))  # the reference passes the voxel size as a float
    size = ctypes.c_uint64()
    call("crvpt_build_fog_volume_sphere", float(radius), c, vs, float(half_width), None, 0, ctypes.byref(size))
    out = np.empty(size.value, dtype=np.uint8)
    call("crvpt_build_fog_volume_sphere", float(radius), c, vs, float(half_width), out.ctypes.data_as(ctypes.c_void_p),
         out.size, ctypes.byref(size))
    return out.tobytes()


def set_sparse_value_type(value_type: int) -> None:
    """Leaf value type of grids set_use_sparse_grid builds from the dense field: 0 FP32, 1 Fp16 (an extension)."""
    call("crvpt_set_sparse_value_type", int(value_type))


def grid_lookup(ijk) -> np.ndarray:
    """Device accessor values at integer coordinates of the current grid (see include/crvpt.h)."""
    a = np.ascontiguousarray(np.asarray(ijk, dtype=np.int32).reshape(-1, 3))
    out = np.empty(a.shape[0], np.float32)
    call("crvpt_grid_lookup_batch", a.shape[0], a.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
    return out


def grid_lookup_world(xyz):
    """Nearest device lookups at NanoVDB world coordinates of the current sparse grid: (values, voxels)."""
    a = np.ascontiguousarray(np.asarray(xyz, dtype=np.float32).reshape(-1, 3))
    out = np.empty(a.shape[0], np.float32)
    ijk = np.empty((a.shape[0], 3), np.int32)
    call("crvpt_grid_lookup_world_batch", a.shape[0], a.ctypes.data_as(ctypes.c_void_p),
         out.ctypes.data_as(ctypes.c_void_p), ijk.ctypes.data_as(ctypes.c_void_p))
    return out, ijk


def get_supervoxel_tables():
    """(edge, grid size, decomposition (min, max) pairs, RRT (mu_c, mu_r_bar) pairs) built on the device."""
    sv = ctypes.c_int()
    g = (ctypes.c_int * 3)()
    call("crvpt_get_supervoxel_tables", ctypes.byref(sv), g, None, None, 0)
    n = g[0] * g[1] * g[2]
    dec = np.empty((n, 2), np.float32)
    rrt = np.empty((n, 2), np.float32)
    call("crvpt_get_supervoxel_tables", ctypes.byref(sv), g, dec.ctypes.data_as(ctypes.c_void_p),
         rrt.ctypes.data_as(ctypes.c_void_p), 2 * n)
    return sv.value, tuple(g), dec, rrt
