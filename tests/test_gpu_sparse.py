"""GPU checks of the sparse (NanoVDB) path: device lookups bit-exact against the reference accessor's golden vectors,
supervoxel tables bit-exact against the oracle's restatement of SuperVoxelGrid.cpp, and sparse renders and
transmittance estimates statistically equal to the dense path and the FP64 optical depth."""
import glob
import math
import os

import numpy as np
import pytest

from _stats import (assert_means_equal, assert_unbiased_binary_rays, assert_unbiased_rays, gpu_image_stats,
                    oracle_image_stats, z_scores_vs_exact)

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = sorted(p for p in glob.glob(os.path.join(REPO, "tests", "golden", "nanovdb_*.npz")) if "fp16" not in p)
GOLDEN16 = sorted(glob.glob(os.path.join(REPO, "tests", "golden", "nanovdb_fp16_*.npz")))


def _cpu(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_device_lookups_bit_exact_vs_reference(vpt, path):
    print(f"case: device NanoVDB lookups {os.path.basename(path)}", flush=True)
    g = np.load(path)
    vpt.set_nanovdb_buffer(g["grid"].tobytes())
    got = vpt.grid_lookup(g["ijk"])
    np.testing.assert_array_equal(got.view(np.uint32), g["values"].view(np.uint32))
    vals, ijk = vpt.grid_lookup_world(g["xyz"])
    np.testing.assert_array_equal(ijk, g["world_ijk"])
    np.testing.assert_array_equal(vals.view(np.uint32), g["world_values"].view(np.uint32))


@pytest.mark.parametrize("path", GOLDEN16, ids=[os.path.basename(p) for p in GOLDEN16])
def test_device_fp16_lookups_bit_exact_vs_reference(vpt, path):
    print(f"case: device Fp16 NanoVDB lookups {os.path.basename(path)}", flush=True)
    g = np.load(path)
    vpt.set_nanovdb_buffer(g["grid"].tobytes())
    got = vpt.grid_lookup(g["ijk"])
    np.testing.assert_array_equal(got.view(np.uint32), g["values"].view(np.uint32))


@pytest.mark.parametrize("mode", [0, 4, 3])
def test_fp16_render_equals_fp32(vpt, oracle, mode):
    """The Fp16 grid quantises each leaf to 1/65535 of its range: renders agree with the FP32 grid's statistically
    (delta, residual ratio and decomposition tracking; residual ratio tracking on Fp16 leaves is the large
    production-frame workload's estimator)."""
    print(f"case: Fp16 vs FP32 sparse render mode={mode}", flush=True)
    d = oracle.fbm_cloud(32, 32, 32, seed=17)
    vpt.set_density_field(d)
    vpt.set_vpt_mode(mode)
    vpt.set_grid_interpolation_type("stochastic")
    vpt.set_extinction_scale(200.0)
    vpt.set_scattering_albedo([0.9, 0.9, 0.9])
    vpt.set_phase_g(0.5)
    vpt.set_use_builtin_environment_map("Default")
    vpt.set_use_lights(False)
    vpt.set_use_sparse_grid(True)
    m32, s32 = _stats(vpt, 64, 32)
    vpt.set_sparse_value_type(1)
    m16, s16 = _stats(vpt, 64, 32)
    vpt.set_sparse_value_type(0)
    vpt.set_use_sparse_grid(False)
    assert_means_equal(m32, s32, m16, s16, f"Fp16 vs FP32 mode {mode}")


def test_coherent_lookups_through_the_memo(vpt, oracle):
    """Queries in x-fastest order over the whole index box plus a random walk: consecutive queries mostly fall in the
    same 8^3 block, so the per-lane lookup memo answers most of them; every answer must equal the voxel written."""
    print("case: coherent NanoVDB lookups 40x36x28", flush=True)
    d = oracle.fbm_cloud(40, 36, 28, seed=13)
    vmin, vmax = float(d.min()), float(d.max())
    norm = ((d - vmin) / (vmax - vmin)).astype(np.float32)
    vpt.set_nanovdb_buffer(vpt.build_nanovdb_from_dense(norm, 0.01))
    k, j, i = np.meshgrid(np.arange(-3, 31), np.arange(-3, 39), np.arange(-3, 43), indexing="ij")
    ijk = np.stack([i.ravel(), j.ravel(), k.ravel()], 1).astype(np.int32)
    rng = np.random.default_rng(3)
    walk = np.cumsum(rng.integers(-1, 2, size=(20000, 3)), 0) + np.array([20, 18, 14])
    walk = np.clip(walk, -2, 44).astype(np.int32)
    q = np.concatenate([ijk, walk])
    got = vpt.grid_lookup(q)
    inside = np.all((q >= 0) & (q < np.array([40, 36, 28])), 1)
    expect = np.zeros(len(q), np.float32)
    expect[inside] = norm[q[inside, 2], q[inside, 1], q[inside, 0]]
    np.testing.assert_array_equal(got.view(np.uint32), expect.view(np.uint32))


def test_tile_blocks_through_the_block_index(vpt):
    """A grid whose interior holds constant 8^3 blocks (stored as tiles, not leaves) next to leaf blocks: lookups go
    through the flat block index and the lane memo (tile blocks memoise the tile value), and queries outside the
    index box descend the tree; every answer equals the voxel written (zero outside)."""
    print("case: NanoVDB tile blocks 48x40x32", flush=True)
    rng = np.random.default_rng(21)
    d = np.zeros((32, 40, 48), np.float32)
    d[8:24, 8:24, 8:32] = 1.0
    d[3:7, 30:35, 20:29] = rng.uniform(0.1, 1.0, (4, 5, 9)).astype(np.float32)
    d[20:29, 0:3, 40:47] = 0.5
    vpt.set_nanovdb_buffer(vpt.build_nanovdb_from_dense(d, 0.01))
    k, j, i = np.meshgrid(np.arange(-9, 41), np.arange(-9, 49), np.arange(-9, 57), indexing="ij")
    q = np.stack([i.ravel(), j.ravel(), k.ravel()], 1).astype(np.int32)
    walk = np.clip(np.cumsum(rng.integers(-2, 3, size=(30000, 3)), 0) + np.array([24, 20, 16]), -12, 60)
    q = np.concatenate([q, walk.astype(np.int32)])
    got = vpt.grid_lookup(q)
    inside = np.all((q >= 0) & (q < np.array([48, 40, 32])), 1)
    expect = np.zeros(len(q), np.float32)
    expect[inside] = d[q[inside, 2], q[inside, 1], q[inside, 0]]
    np.testing.assert_array_equal(got.view(np.uint32), expect.view(np.uint32))


def test_grid_beyond_4_gib(vpt):
    """A NanoVDB grid of 5.3 GiB (1100^3 voxels, all active): leaves beyond the first 4 GiB are read through 64-bit
    offsets, via the block index and via tree descents outside it; every lookup equals the voxel written."""
    import ctypes

    print("case: 1100^3 dense-active NanoVDB grid (5.3 GiB)", flush=True)
    n = 1100
    vpt.generate_fbm_cloud(n, n, n, seed=3, octaves=5, edge0=-1.0, edge1=2.0)
    vpt.set_use_sparse_grid(True)
    rng = np.random.default_rng(5)
    q = rng.integers(-3, n + 3, size=(500_000, 3)).astype(np.int32)
    top = np.stack(np.meshgrid(np.arange(n), np.arange(n - 8, n), np.arange(n - 4, n), indexing="ij"), -1)
    q = np.concatenate([q, top.reshape(-1, 3).astype(np.int32)])
    got = vpt.grid_lookup(q)  # builds the grid (prepareScene) before the size query below
    size = ctypes.c_uint64()
    vpt.call("crvpt_get_nanovdb_buffer", None, 0, ctypes.byref(size))
    assert size.value > 2 ** 32 + 2 ** 30
    d = vpt.get_density_field((n, n, n), normalized=True)
    inside = np.all((q >= 0) & (q < n), 1)
    expect = np.zeros(len(q), np.float32)
    expect[inside] = d[q[inside, 2], q[inside, 1], q[inside, 0]]
    del d
    np.testing.assert_array_equal(got.view(np.uint32), expect.view(np.uint32))
    # the single-frame path kernels' 64-bit variant (grids of 4 GiB and more) against the dense field of the same
    # volume: delta tracking, stochastic filtering, a thin medium so paths stay short
    print("case: 1100^3 render, sparse 64-bit path kernels vs dense", flush=True)
    vpt.set_vpt_mode(0)
    vpt.set_grid_interpolation_type("stochastic")
    vpt.set_extinction_base([1.0, 1.0, 1.0])
    vpt.set_extinction_scale(6.0)
    vpt.set_scattering_albedo([0.9, 0.9, 0.9])
    vpt.set_phase_g(0.5)
    vpt.set_use_builtin_environment_map("Default")
    vpt.set_use_lights(False)
    vpt.set_seed_offset(0)
    m_s, s_s, _ = gpu_image_stats(vpt, 256, 24)
    vpt.set_use_sparse_grid(False)
    m_d, s_d, _ = gpu_image_stats(vpt, 256, 24)
    assert_means_equal(m_s, s_s, m_d, s_d, "1100^3 sparse (64-bit offsets) vs dense")
    assert m_d.min() > 0.01


@pytest.mark.parametrize("interp", ["nearest", "stochastic"])
@pytest.mark.parametrize("sparse", [False, True])
@pytest.mark.parametrize("sv", [8, 16])
def test_supervoxel_tables_bit_exact(vpt, oracle, interp, sparse, sv):
    print(f"case: supervoxel tables interp={interp} sparse={sparse} edge={sv}", flush=True)
    d = oracle.fbm_cloud(44, 36, 28, seed=12)
    vpt.set_density_field(d)
    vpt.set_super_voxel_size(sv)
    vpt.set_grid_interpolation_type(interp)
    vpt.set_extinction_scale(300.0)
    vpt.set_extinction_base([1.0, 1.0, 1.0])
    vpt.set_use_sparse_grid(sparse)
    sv, g, dec, rrt = vpt.get_supervoxel_tables()
    vmin, vmax = float(d.min()), float(d.max())
    norm = ((d - vmin) / (vmax - vmin)).astype(np.float32)
    if sparse:
        nz_, ny_, nx_ = np.nonzero(norm != 0)
        crop = norm[nz_.min():nz_.max() + 1, ny_.min():ny_.max() + 1, nx_.min():nx_.max() + 1]
        sc = oracle.Scene(crop, interpolation=interp, extinction=300.0, voxel_range=(0.0, 1.0), super_voxel_size=sv)
    else:
        sc = oracle.Scene(d, interpolation=interp, extinction=300.0, super_voxel_size=sv)
    osv, og, orrt, odec, _ = sc.supervoxel_tables()
    vpt.set_use_sparse_grid(False)
    vpt.set_super_voxel_size(8)
    assert (sv, g) == (osv, og)
    np.testing.assert_array_equal(dec.view(np.uint32), odec.view(np.uint32))
    np.testing.assert_array_equal(rrt.view(np.uint32), orrt.view(np.uint32))


def _stats(vpt, frames, res):
    """Image mean and its standard error from one render of 4 x frames samples per pixel (per-pixel variances)."""
    m, se, _ = gpu_image_stats(vpt, 4 * frames, res)
    return m, se


def test_nanovdb_only_volume_renders_dense(vpt, oracle):
    """CloudData::getDenseDensityField (CloudData.cpp:1173-1253): a volume loaded only as NanoVDB and rendered with
    set_use_sparse_grid(False) goes through a dense field decoded from the grid over its index bounding box, placed
    in the sparse grid's box. The decoded field equals the grid's source voxels bit for bit, and its renders agree
    with the sparse renders within 3 sigma (before, this switch left the renderer reading a stale dense buffer)."""
    print("case: NanoVDB-only volume rendered dense", flush=True)
    d = oracle.fbm_cloud(24, 20, 28, seed=5)
    norm = ((d - d.min()) / (d.max() - d.min())).astype(np.float32)
    nz_idx = np.argwhere(norm != 0.0)
    (z0, y0, x0), (z1, y1, x1) = nz_idx.min(0), nz_idx.max(0)
    vpt.set_vpt_mode(0)
    vpt.set_grid_interpolation_type("nearest")
    vpt.set_extinction_scale(150.0)
    vpt.set_scattering_albedo([0.9, 0.9, 0.9])
    vpt.set_phase_g(0.5)
    vpt.set_use_builtin_environment_map("Default")
    vpt.set_use_lights(False)
    vpt.set_nanovdb_buffer(vpt.build_nanovdb_from_dense(norm, 0.01))
    box_sparse = vpt.get_render_bounding_box()
    m_sparse, s_sparse = _stats(vpt, 32, 32)
    vpt.set_use_sparse_grid(False)
    assert vpt.get_volume_voxel_size() == [int(z1 - z0 + 1), int(y1 - y0 + 1), int(x1 - x0 + 1)]
    dec = vpt.get_density_field((z1 - z0 + 1, y1 - y0 + 1, x1 - x0 + 1))
    np.testing.assert_array_equal(dec, norm[z0:z1 + 1, y0:y1 + 1, x0:x1 + 1])
    m_dense, s_dense = _stats(vpt, 32, 32)
    assert vpt.get_render_bounding_box() == box_sparse
    assert_means_equal(m_sparse, s_sparse, m_dense, s_dense, "NanoVDB-only volume sparse vs dense")
    vpt.set_use_sparse_grid(True)
    m_back, s_back = _stats(vpt, 32, 32)
    np.testing.assert_array_equal(m_back, m_sparse)  # same seeds, same grid: identical images


@pytest.mark.parametrize("name,mode", [("delta", 0), ("next_event", 5), ("residual_ratio", 4)])
@pytest.mark.parametrize("interp", ["trilinear", "stochastic"])
def test_sparse_render_vs_oracle(vpt, oracle, name, mode, interp):
    """NanoVDB renders (the path kernels' 32-bit accessor; trilinear: the 8-tap variant, VptUtils.glsl:515-580)
    against the oracle on the grid's values: raw NanoVDB values, no min-max normalisation (VptUtils.glsl:523),
    within the grid's active index box."""
    print(f"case: sparse render vs oracle {name} {interp}", flush=True)
    d = oracle.fbm_cloud(32, 28, 24, seed=19)
    vmin, vmax = float(d.min()), float(d.max())
    norm = ((d - vmin) / (vmax - vmin)).astype(np.float32)
    vpt.set_density_field(d)
    vpt.set_vpt_mode(mode)
    vpt.set_grid_interpolation_type(interp)
    vpt.set_extinction_scale(150.0)
    vpt.set_extinction_base([1.0, 1.0, 1.0])
    vpt.set_scattering_albedo([0.9, 0.9, 0.9])
    vpt.set_phase_g(0.5)
    vpt.set_use_builtin_environment_map("Default")
    vpt.set_use_lights(False)
    vpt.set_seed_offset(0)
    vpt.set_use_sparse_grid(True)
    box = vpt.get_render_bounding_box()
    try:
        m_gpu, s_gpu, _ = gpu_image_stats(vpt, 256, 32)
    finally:
        vpt.set_use_sparse_grid(False)
    nz_, ny_, nx_ = np.nonzero(norm != 0)
    crop = np.ascontiguousarray(norm[nz_.min():nz_.max() + 1, ny_.min():ny_.max() + 1, nx_.min():nx_.max() + 1])
    sc = oracle.Scene(crop, mode=name, interpolation=interp, extinction=150.0, g=0.5, width=32, height=32,
                      voxel_range=(0.0, 1.0),
                      box=(np.array(box[0::2], np.float32), np.array(box[1::2], np.float32)))
    m_ref, s_ref, _ = oracle_image_stats(sc, 64)
    assert_means_equal(m_gpu, s_gpu, m_ref, s_ref, f"sparse {name}/{interp}")


@pytest.mark.parametrize("mode", [0, 2, 5, 3, 4])
def test_sparse_render_equals_dense(vpt, oracle, mode):
    """TestVolumetricPathTracing.cpp:187-197 (dense vs sparse grid types) for each estimator."""
    print(f"case: sparse vs dense render mode={mode}", flush=True)
    d = oracle.fbm_cloud(32, 32, 32, seed=13)
    vpt.set_density_field(d)
    vpt.set_vpt_mode(mode)
    vpt.set_grid_interpolation_type("nearest")
    vpt.set_extinction_scale(200.0)
    vpt.set_scattering_albedo([0.9, 0.9, 0.9])
    vpt.set_phase_g(0.5)
    vpt.set_use_builtin_environment_map("Default")
    vpt.set_use_lights(False)
    vpt.set_use_sparse_grid(False)
    md, sd = _stats(vpt, 64, 32)
    vpt.set_use_sparse_grid(True)
    ms, ss = _stats(vpt, 64, 32)
    vpt.set_use_sparse_grid(False)
    assert_means_equal(md, sd, ms, ss, f"sparse vs dense mode {mode}")


@pytest.mark.parametrize("estimator", ["delta", "ratio", "residual_ratio", "decomposition"])
@pytest.mark.parametrize("interp", ["nearest", "stochastic", "trilinear"])
def test_sparse_transmittance_vs_fp64(vpt, oracle, estimator, interp):
    print(f"case: sparse transmittance {estimator} {interp}", flush=True)
    d = oracle.fbm_cloud(40, 40, 40, seed=14)
    vpt.set_density_field(d)
    vpt.set_grid_interpolation_type(interp)
    vpt.set_extinction_scale(40.0)
    vpt.set_extinction_base([1.0, 1.0, 1.0])
    vpt.set_use_sparse_grid(True)
    box = vpt.get_render_bounding_box()
    bmin = np.array(box[0::2], np.float32)
    bmax = np.array(box[1::2], np.float32)
    rng = np.random.default_rng(15)
    n_rays, n = 512, 8192
    c = 0.5 * (bmin + bmax)
    half = 0.5 * (bmax - bmin)
    o = rng.normal(size=(n_rays, 3))
    o = (c + o / np.linalg.norm(o, axis=1, keepdims=True) * 1.5 * np.linalg.norm(half)).astype(np.float32)
    tgt = rng.uniform(bmin, bmax, (n_rays, 3))
    w = (tgt - o) / np.linalg.norm(tgt - o, axis=1, keepdims=True)
    w = w.astype(np.float32)
    mean, m2 = vpt.transmittance_batch(estimator, o, w, n_samples=n, seed=9)
    vpt.set_use_sparse_grid(False)
    vmin, vmax = float(d.min()), float(d.max())
    norm = ((d - vmin) / (vmax - vmin)).astype(np.float32)
    nz_, ny_, nx_ = np.nonzero(norm != 0)
    crop = norm[nz_.min():nz_.max() + 1, ny_.min():ny_.max() + 1, nx_.min():nx_.max() + 1]
    sc = oracle.Scene(crop, interpolation=interp, extinction=40.0, voxel_range=(0.0, 1.0), box=(bmin, bmax))
    exact = np.exp(-sc.optical_depth(o, w))
    if estimator in ("delta", "decomposition"):  # tracking estimators: every sample is 0 or 1
        assert_unbiased_binary_rays(mean, m2, exact, n, f"sparse transmittance {estimator}/{interp}")
    else:
        assert_unbiased_rays(z_scores_vs_exact(mean, m2, exact, n), f"sparse transmittance {estimator}/{interp}")
