"""GPU parity on the benchmark workloads' own scenes (the volumes, estimators and settings bench.py times).

* C1 dense 128^3 fBm, next-event tracking with a constant white environment and a directional sun, g 0.85,
  extinction 1024, at most 64 scattering events: image means at 64 x 64 against the oracle, and a converged crop.
* C2 NanoVDB 512^3 fBm (about 12 % active, 16^3 supervoxels), the four timed estimators x g in {0, 0.5, 0.85},
  stochastic filtering, at most 128 scattering events: image means at 64 x 64 against the oracle on the grid's
  values (the dense-equivalent crop of the active index box, raw values, the grid's box), and one converged crop
  per estimator.
* C3 the 64 requests of the batched queue at spp / 16 in one batch launch: every request equals its render_frame.
* C5 NanoVDB Fp16 2048^3 anvil fBm: device lookups at 10^6 voxels bit-exact against the reference PNanoVDB
  accessor reading the same grid (oracle/_ref/nvdb_ref), and a residual-ratio render of the production view at
  192 x 108 against the same view on the FP32 grid.

Image means: z-test at the family-wise 3-sigma level of the three channels (tests/_stats.py). Converged crops: per-pixel
z-scores from both renders' per-pixel variances, unbiased on average per channel (tests/_stats.py assert_converged).
"""
import math
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from _stats import assert_converged, assert_means_equal, gpu_image_stats, oracle_image_stats

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
import bench  # noqa: E402  (the workloads' scene definitions)

TOOL = os.path.join(REPO, "oracle", "_ref", "nvdb_ref")
THREADS = 16


def _torch():
    import torch

    return torch


def _box(vpt):
    b = vpt.get_render_bounding_box()
    return np.array(b[0::2], np.float32), np.array(b[1::2], np.float32)


def _converged_crop(vpt, oracle, scene_kwargs, density, target, what, res=6, gpu_spp=65536, ref_spp=4096):
    """A res x res crop around `target` (fovy 2 atan(0.5 / 8): one eighth of the default view) from the current
    camera position, rendered to convergence on the GPU and by the oracle, compared pixel by pixel
    (tests/_stats.py assert_converged)."""
    torch = _torch()
    fovy = 2.0 * math.atan(0.5 / 8.0)
    pos = vpt.get_camera_position()
    # set_camera_target looks from the position last given to set_camera_position (the module renderer's member,
    # VolumetricPathTracingModuleRenderer.cpp:308-319), so the position is set first
    vpt.set_camera_position(list(pos))
    vpt.set_camera_target(list(float(v) for v in target))
    vpt.set_camera_fovy(math.degrees(fovy))
    vpt.set_use_feature_maps(["Radiance Variance"])
    try:
        t = torch.zeros((4, res, res), dtype=torch.float32, device="cuda")
        gpu = vpt.render_frame(t, gpu_spp).cpu().numpy()[:3].transpose(1, 2, 0).astype(np.float64)
        var_gpu = vpt.get_feature_map(t, 21).cpu().numpy()[:3].transpose(1, 2, 0).astype(np.float64)
    finally:
        vpt.set_use_feature_maps([])
        vpt.set_camera_fovy(math.degrees(2.0 * math.atan(0.5)))
        vpt.set_camera_target([pos[0], pos[1], pos[2] - 1.0])
    cam = oracle.camera_matrices(pos=tuple(pos), target=tuple(float(v) for v in target), fovy=fovy, aspect=1.0)
    sc = oracle.Scene(density, camera=cam, width=res, height=res, **scene_kwargs)
    ref, sq = sc.render_mean(ref_spp, threads=THREADS)
    var_ref = np.maximum(sq / ref_spp - ref ** 2, 0.0)
    assert ref.mean() > 0.01, "the crop must see the cloud"
    return assert_converged(gpu, var_gpu, gpu_spp, ref, var_ref, ref_spp, what)


def _centre_of_mass(density, bmin, bmax):
    nz, ny, nx = density.shape
    w = density.astype(np.float64)
    tot = w.sum()
    c = [float((w.sum(axis=ax) * (np.arange(n) + 0.5)).sum() / tot / n)
         for ax, n in (((0, 1), nx), ((0, 2), ny), ((1, 2), nz))]
    return bmin + np.array(c, np.float32) * (bmax - bmin)


# --------------------------------------------------------------------------------------------------------------
# C1
# --------------------------------------------------------------------------------------------------------------
def _c1_oracle_kwargs():
    sun = np.array([0.5826, 0.7660, 0.2717])
    sun = 100.0 * sun / np.linalg.norm(sun)
    from oracle import oracle as O

    L = O.make_light(light_type=2, light_space=0, position=tuple(sun), intensity=2.6,
                     color=(1.0, 0.961538462, 0.884615385))
    return dict(mode="next_event", interpolation="stochastic", extinction=1024.0, albedo=0.9, g=0.85,
                env_image=np.ones((3, 16, 32), np.float32), lights=[L], max_grid_extent=0.5,
                max_bounces=bench.C1["max_bounces"])


def test_c1_dense128_next_event_vs_oracle(vpt, oracle):
    print("case: C1 dense 128^3 next-event env + sun, 64x64", flush=True)
    torch = _torch()
    bench.setup_c1(vpt, torch)
    vpt.set_seed_offset(0)
    try:
        m_gpu, s_gpu, _ = gpu_image_stats(vpt, 4096, 64)
        d = oracle.fbm_cloud(128, 128, 128, seed=0)
        np.testing.assert_array_equal(vpt.get_density_field((128, 128, 128)), d)
        sc = oracle.Scene(d, width=64, height=64, **_c1_oracle_kwargs())
        m_ref, s_ref, _ = oracle_image_stats(sc, 256, threads=THREADS)
        assert_means_equal(m_gpu, s_gpu, m_ref, s_ref, "C1 next-event env + sun")
        bmin, bmax = _box(vpt)
        _converged_crop(vpt, oracle, _c1_oracle_kwargs(), d, _centre_of_mass(d, bmin, bmax), "C1 converged crop")
    finally:
        vpt.set_max_grid_extent(0.25)
        vpt.set_use_lights(False)
        vpt.set_max_bounces(0)


# --------------------------------------------------------------------------------------------------------------
# C3
# --------------------------------------------------------------------------------------------------------------
def test_c3_batch_requests_match_render_frame(vpt):
    """The batched-queue workload's 64 requests (bench.c3_requests) at spp / 16, one sample per unit: each request's
    image equals render_frame of the same scene up to the rounding of the two kernels (tests/test_gpu_batch.py)."""
    print("case: C3 64 requests at spp/16 vs render_frame", flush=True)
    torch = _torch()
    div = 16
    reqs = bench.setup_c3(vpt, torch, spp_divisor=div)
    vpt.batch_set_launch_config(0, 256, 16, 1)
    outs = [o.cpu().numpy() for o in vpt.batch_render(0)]
    torch.cuda.synchronize()
    st = vpt.batch_get_stats()
    assert st["num_requests"] == len(reqs) == 64
    order = [q for kind in range(len(bench.C3["volumes"])) for q in reqs if q["kind"] == kind]
    assert st["num_paths"] == sum(q["side"] ** 2 * bench.c3_spp(q, div) for q in reqs)
    bench.c3_settings(vpt)
    worst = 0.0
    k = 0
    for kind in range(len(bench.C3["volumes"])):
        mine = [q for q in order if q["kind"] == kind]
        if not mine:
            continue
        bench.c3_volume(vpt, kind)
        for q in mine:
            bench.c3_view(vpt, q)
            ref = vpt.render_frame(torch.zeros((3, q["side"], q["side"]), dtype=torch.float32, device="cuda"),
                                   bench.c3_spp(q, div)).cpu().numpy()
            got = outs[k]
            k += 1
            assert np.isfinite(got).all(), f"request {q['index']}"
            close = np.isclose(got, ref, rtol=2e-6, atol=2e-7)
            worst = max(worst, 1.0 - close.mean())
            assert close.mean() >= 0.998, f"request {q['index']}: {int((~close).sum())} of {close.size} differ"
            assert abs(float(got.mean()) - float(ref.mean())) <= 1e-6 * max(abs(float(ref.mean())), 1e-3)
    print(f"C3: largest fraction of values beyond 2e-6 relative: {worst:.2e}", flush=True)
    vpt.batch_clear()
    vpt.set_use_lights(False)
    vpt.set_max_bounces(0)
    vpt.set_seed_offset(0)
    vpt.set_camera_position([0.0, 0.0, 0.8])
    vpt.set_camera_target([0.0, 0.0, 0.0])


# --------------------------------------------------------------------------------------------------------------
# C2
# --------------------------------------------------------------------------------------------------------------
C2_NAMES = {0: "delta", 2: "ratio", 4: "residual_ratio", 3: "decomposition"}


@pytest.fixture(scope="module")
def c2(vpt):
    print("case: C2 volume 512^3 sparse and its dense-equivalent crop", flush=True)
    bench.setup_c2(vpt)
    g = bench.C2["grid"]
    d = vpt.get_density_field((g, g, g), normalized=True)
    nzm = d != 0
    zs, ys, xs = (np.nonzero(nzm.any(axis=ax))[0] for ax in ((1, 2), (0, 2), (0, 1)))
    del nzm
    crop = np.ascontiguousarray(d[zs[0]:zs[-1] + 1, ys[0]:ys[-1] + 1, xs[0]:xs[-1] + 1])
    del d
    vpt.render_frame(_torch().zeros((3, 8, 8), dtype=_torch().float32, device="cuda"), 1)  # builds the grid
    return {"crop": crop, "box": _box(vpt)}


def _c2_kwargs(c2, mode, g):
    return dict(mode=C2_NAMES[mode], interpolation="stochastic", extinction=1024.0, albedo=0.9, g=g,
                voxel_range=(0.0, 1.0), box=c2["box"], super_voxel_size=16, max_bounces=bench.C2["max_bounces"])


@pytest.mark.parametrize("g", bench.C2["gs"])
@pytest.mark.parametrize("mode", [m for m, _ in bench.C2["modes"]])
def test_c2_sparse512_vs_oracle(vpt, oracle, c2, mode, g):
    print(f"case: C2 512^3 sparse {C2_NAMES[mode]} g={g}, 64x64", flush=True)
    vpt.set_use_sparse_grid(True)
    vpt.set_vpt_mode(mode)
    vpt.set_phase_g(g)
    vpt.set_seed_offset(0)
    m_gpu, s_gpu, _ = gpu_image_stats(vpt, 1024, 64)
    sc = oracle.Scene(c2["crop"], width=64, height=64, **_c2_kwargs(c2, mode, g))
    m_ref, s_ref, _ = oracle_image_stats(sc, 32, threads=THREADS)
    assert_means_equal(m_gpu, s_gpu, m_ref, s_ref, f"C2 {C2_NAMES[mode]} g={g}")


@pytest.mark.parametrize("mode", [m for m, _ in bench.C2["modes"]])
def test_c2_sparse512_converged_crop(vpt, oracle, c2, mode):
    print(f"case: C2 converged crop {C2_NAMES[mode]}", flush=True)
    vpt.set_use_sparse_grid(True)
    vpt.set_vpt_mode(mode)
    vpt.set_phase_g(0.5)
    vpt.set_seed_offset(0)
    bmin, bmax = c2["box"]
    target = _centre_of_mass(c2["crop"], bmin, bmax)
    _converged_crop(vpt, oracle, _c2_kwargs(c2, mode, 0.5), c2["crop"], target, f"C2 {C2_NAMES[mode]} converged crop")


def test_c2_cleanup(vpt):
    vpt.set_use_sparse_grid(False)
    vpt.set_launch_config(0, 0, 0)
    vpt.set_super_voxel_size(8)
    vpt.set_max_bounces(0)


# --------------------------------------------------------------------------------------------------------------
# C5
# --------------------------------------------------------------------------------------------------------------
@pytest.mark.skipif(not os.path.exists(TOOL), reason="reference PNanoVDB tool not built (oracle/_ref)")
def test_c5_fp16_2048_lookups_and_render(vpt):
    """The production frame's Fp16 grid: 10^6 device lookups (random voxels of the index box, plus the 512 voxels
    of each of 64 leaves) equal the reference accessor's on the same grid bit for bit; the residual-ratio render of
    the first production view at 192 x 108 matches the FP32 grid's render of the same view."""
    print("case: C5 2048^3 Fp16 grid lookups and render", flush=True)
    torch = _torch()
    bench.setup_c5(vpt)
    bench._c5_camera(vpt, 0)
    vpt.render_frame(torch.zeros((3, 8, 8), dtype=torch.float32, device="cuda"), 1)  # builds the Fp16 grid
    buf = np.frombuffer(vpt.get_nanovdb_buffer(), np.uint8)
    assert int.from_bytes(buf[636:640].tobytes(), "little") == 15  # GridType::Fp16
    print(f"C5 Fp16 grid: {buf.size / 2 ** 30:.3f} GiB", flush=True)
    rng = np.random.default_rng(55)
    lo = np.frombuffer(buf[672 + 64:672 + 64 + 12].tobytes(), np.int32)
    hi = np.frombuffer(buf[672 + 64 + 12:672 + 64 + 24].tobytes(), np.int32)
    ijk = rng.integers(lo, hi + 1, size=(1_000_000 - 64 * 512, 3)).astype(np.int32)
    # whole leaves around random active voxels
    leaves = ijk[:64] & ~7
    k, j, i = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    cube = np.stack([i.ravel(), j.ravel(), k.ravel()], 1)
    ijk = np.concatenate([ijk, (leaves[:, None, :] + cube[None]).reshape(-1, 3)]).astype(np.int32)
    got = vpt.grid_lookup(ijk)
    with tempfile.TemporaryDirectory() as tmp:
        g, q, o = (os.path.join(tmp, n) for n in ("g.bin", "ijk.i32", "v.f32"))
        buf.tofile(g)
        ijk.tofile(q)
        subprocess.run([TOOL, "ijk16", g, q, o], check=True)
        ref = np.fromfile(o, np.float32)
    del buf
    active = float(np.count_nonzero(ref)) / ref.size
    print(f"C5 lookups: {ijk.shape[0]} voxels, {active:.3f} nonzero", flush=True)
    assert active > 0.01
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    m16, s16, _ = gpu_image_stats(vpt, 256, (192, 108))
    vpt.set_sparse_value_type(0)  # rebuilds the FP32 grid of the same field
    try:
        m32, s32, _ = gpu_image_stats(vpt, 256, (192, 108))
    finally:
        vpt.set_sparse_value_type(1)
    assert_means_equal(m16, s16, m32, s32, "C5 Fp16 vs FP32 residual ratio tracking")
    vpt.set_use_sparse_grid(False)
    vpt.set_sparse_value_type(0)
    vpt.set_max_bounces(0)
    vpt.set_super_voxel_size(8)
