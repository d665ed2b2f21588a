"""GPU checks of the module's scene setup: the default scene (Module.cpp:793-795), the foreground map output
(set_output_foreground_map, Clouds.glsl:181-191) and the file loaders (load_cloud_file for .nvdb and .xyz,
CloudData.cpp:236-316, 1360-1365).

The default-scene tests must run first in this module: the module's renderer is freshly initialised, so it still
holds the default scene.
"""
import os
import struct

import numpy as np
import pytest

from _stats import assert_means_equal, gpu_image_stats, oracle_image_stats

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _torch():
    import torch

    return torch


def _render(vpt, frames, res, channels=3):
    torch = _torch()
    return vpt.render_frame(torch.zeros((channels, res, res), dtype=torch.float32, device="cuda"), frames)


def test_default_scene_is_the_reference_fog_sphere(vpt):
    """initialize() loads createFogVolumeSphere<float>(0.25, 0, 0.01): the grid equals the reference's byte for byte,
    device lookups equal the reference accessor's (including the -0.0 voxels on the sphere surface), and the dense
    field the pass renders (decoded from the grid over its index box) holds the same values."""
    print("case: default scene grid and lookups", flush=True)
    g = np.load(os.path.join(GOLDEN, "fog_sphere_default.npz"))
    buf = np.frombuffer(vpt.get_nanovdb_buffer(), np.uint8)
    np.testing.assert_array_equal(buf, g["grid"])
    vpt.set_use_sparse_grid(True)  # grid_lookup reads the active representation; the default one is dense
    try:
        got = vpt.grid_lookup(g["ijk"])
    finally:
        vpt.set_use_sparse_grid(False)
    np.testing.assert_array_equal(got.view(np.uint32), g["values"].view(np.uint32))
    assert vpt.get_volume_voxel_size() == [51, 51, 51]  # index box [-25, 25]^3
    dense = vpt.get_density_field((51, 51, 51))
    ijk = g["ijk"]
    inside = np.all(np.abs(ijk) <= 25, 1)
    q = ijk[inside] + 25
    np.testing.assert_array_equal(dense[q[:, 2], q[:, 1], q[:, 0]].view(np.uint32),
                                  g["values"][inside].view(np.uint32))
    box = vpt.get_render_bounding_box()
    np.testing.assert_allclose(box, [-0.25, 0.25] * 3, rtol=1e-6)


def test_default_scene_render_vs_oracle(vpt, oracle):
    """The default scene at the default settings (delta tracking, stochastic filtering, extinction 1024, albedo 0.9,
    g 0.5, default sky, the default camera) against the oracle on the same dense field and box."""
    print("case: default scene render vs oracle", flush=True)
    dense = vpt.get_density_field((51, 51, 51))
    box = vpt.get_render_bounding_box()
    m_gpu, s_gpu, _ = gpu_image_stats(vpt, 256, 32)
    sc = oracle.Scene(dense, mode="delta", width=32, height=32,
                      box=(np.array(box[0::2], np.float32), np.array(box[1::2], np.float32)))
    m_ref, s_ref, _ = oracle_image_stats(sc, 64)
    assert_means_equal(m_gpu, s_gpu, m_ref, s_ref, "default scene delta tracking")
    assert m_ref.min() > 0.05  # the sphere and the sky are both in view


def _fbm_scene(vpt, oracle, mode, ext=96.0):
    d = oracle.fbm_cloud(24, 20, 16, seed=5)
    vpt.set_density_field(d)
    vpt.set_vpt_mode(mode)
    vpt.set_grid_interpolation_type("stochastic")
    vpt.set_extinction_scale(ext)
    vpt.set_extinction_base([1.0, 1.0, 1.0])
    vpt.set_scattering_albedo([0.9, 0.9, 0.9])
    vpt.set_phase_g(0.5)
    vpt.set_use_builtin_environment_map("Default")
    vpt.set_use_lights(False)
    vpt.set_seed_offset(0)
    return d


@pytest.mark.parametrize("name,mode", [("delta", 0), ("ratio", 2), ("next_event", 5), ("delta_spectral", 1)])
def test_foreground_map_is_the_cloud_only_store(vpt, oracle, name, mode):
    """With set_output_foreground_map(True) the result is the OUTPUT_FOREGROUND_MAP store (Clouds.glsl:181-191):
    the running mean of (radiance, 1) over samples whose first event has a value, else (0, 0, 0, 0). That is the
    Cloud Only map's formula (Clouds.glsl:151-160), so in one render the two agree pixel by pixel (up to the
    different float summation of the two accumulations); the coverage (alpha) matches the oracle's."""
    print(f"case: foreground map {name}", flush=True)
    torch = _torch()
    d = _fbm_scene(vpt, oracle, mode)
    res, spp = 32, 128
    vpt.set_output_foreground_map(True)
    vpt.set_use_feature_maps(["Cloud Only"])
    try:
        t = torch.zeros((4, res, res), dtype=torch.float32, device="cuda")
        fg = vpt.render_frame(t, spp).cpu().numpy()
        co = vpt.get_feature_map_from_string(t, "Cloud Only").cpu().numpy()
    finally:
        vpt.set_use_feature_maps([])
        vpt.set_output_foreground_map(False)
    np.testing.assert_allclose(fg, co, rtol=2e-5, atol=2e-6)
    cov = fg[3]
    assert 0.05 < cov.mean() < 0.95
    sc = oracle.Scene(d, mode=name, extinction=96.0, g=0.5, width=res, height=res)
    ref = sc.render_features(spp, names=["cloud_only"])["cloud_only"]
    cov_ref = ref[..., 3].astype(np.float64)
    # coverage of a pixel is a Bernoulli mean: variance c (1 - c) / spp per pixel on each side
    se = np.sqrt((cov_ref * (1.0 - cov_ref)).sum() / spp) / (res * res)
    assert_means_equal(cov.mean(), se, cov_ref.mean(), se, f"{name} foreground coverage")


@pytest.mark.parametrize("mode", [3, 4])
def test_foreground_map_without_first_events_is_zero(vpt, oracle, mode):
    """Decomposition and residual ratio tracking never record a first event (DecompositionTracking.glsl,
    ResidualRatioTracking.glsl), so their foreground map stores (0, 0, 0, 0) for every sample."""
    print(f"case: foreground map mode {mode}", flush=True)
    _fbm_scene(vpt, oracle, mode)
    vpt.set_output_foreground_map(True)
    try:
        fg = _render(vpt, 16, 24, channels=4).cpu().numpy()
    finally:
        vpt.set_output_foreground_map(False)
    assert np.all(fg == 0.0)
    plain = _render(vpt, 16, 24, channels=4).cpu().numpy()
    assert plain[:3].mean() > 0.0 and np.all(plain[3] == 1.0)


def test_foreground_map_setting_resets_accumulation(vpt, oracle):
    print("case: foreground map resets accumulation", flush=True)
    _fbm_scene(vpt, oracle, 0)
    a = _render(vpt, 8, 16, channels=4).cpu().numpy()
    vpt.set_output_foreground_map(True)
    vpt.set_output_foreground_map(False)
    b = _render(vpt, 8, 16, channels=4).cpu().numpy()
    np.testing.assert_array_equal(a, b)


def test_foreground_map_rejected_by_the_batch_queue(vpt, oracle):
    torch = _torch()
    _fbm_scene(vpt, oracle, 0)
    vpt.batch_clear()
    vpt.set_output_foreground_map(True)
    try:
        with pytest.raises(RuntimeError):
            vpt.batch_add_request(torch.zeros((3, 8, 8), dtype=torch.float32, device="cuda"), 4)
    finally:
        vpt.set_output_foreground_map(False)
        vpt.batch_clear()


def test_nvdb_file_load(vpt, oracle, tmp_path):
    """load_cloud_file on a .nvdb file written by the reference's io::writeGrid: the renderer holds the file's grid
    byte for byte; a grid below 1024^3 index entries keeps the dense path (VolumetricPathTracingPass.cpp:780-795),
    rendering the field decoded from the grid, exactly as the same grid given as a buffer and switched to dense."""
    print("case: .nvdb load", flush=True)
    f = np.load(os.path.join(GOLDEN, "nvdb_file_block20.npz"))["nvdb"]
    grid = np.load(os.path.join(GOLDEN, "nanovdb_block20.npz"))["grid"]
    p = str(tmp_path / "block20.nvdb")
    f.tofile(p)
    _fbm_scene(vpt, oracle, 0, ext=64.0)
    vpt.load_cloud_file(p)
    np.testing.assert_array_equal(np.frombuffer(vpt.get_nanovdb_buffer(), np.uint8), grid)
    a = _render(vpt, 16, 24).cpu().numpy()
    vpt.set_nanovdb_buffer(grid.tobytes())
    vpt.set_use_sparse_grid(False)
    b = _render(vpt, 16, 24).cpu().numpy()
    np.testing.assert_array_equal(a, b)


def test_xyz_file_load(vpt, oracle, tmp_path):
    """load_cloud_file on an .xyz file: the field is transposed from the file's z-fastest order and min-max
    normalised (CloudData.cpp:236-316), and renders exactly as the same normalised field set directly."""
    print("case: .xyz load", flush=True)
    d = oracle.fbm_cloud(24, 20, 16, seed=9) * np.float32(3.0) + np.float32(0.25)
    nz, ny, nx = d.shape
    p = str(tmp_path / "f.xyz")
    with open(p, "wb") as fh:
        fh.write(struct.pack("<3I3d", nx, ny, nz, 1.0, 1.0, 1.0))
        fh.write(np.ascontiguousarray(d.transpose(2, 1, 0)).astype("<f4").tobytes())
    _fbm_scene(vpt, oracle, 0, ext=64.0)
    vpt.load_cloud_file(p)
    norm = (d / d.max()).astype(np.float32)  # minimum seeded with 0
    np.testing.assert_array_equal(vpt.get_density_field((nz, ny, nx)), norm)
    a = _render(vpt, 16, 24).cpu().numpy()
    box_a = vpt.get_render_bounding_box()
    vpt.set_density_field(norm)
    b = _render(vpt, 16, 24).cpu().numpy()
    assert vpt.get_render_bounding_box() == box_a
    np.testing.assert_array_equal(a, b)
