"""CPU checks that pin the oracle (oracle/vpt_oracle.cpp) before it checks the GPU path.

Each test fixes a property the mathematics or the reference's own tests determine, chosen so that a plausible
restatement mistake (wrong constant, flipped axis, dropped term) fails it:
  * hybrid Taus: seed 0 leaves the three Tausworthe words at 0, so the stream is the bare LCG (1664525,
    1013904223) after the seed % 7 + 2 = 2 warm-up draws, scaled by 2^-32 (VptUtils.glsl:46-65);
  * camera: centre ray along -z, top image row looks up, edge rays at fovy/2 (VolumetricPathTracingModuleRenderer.cpp:
    124-131);
  * FP64 optical depth: homogeneous cube chord lengths (nearest) and the half-voxel trilinear border ramp;
    brute-force quadrature on random rays through a random grid;
  * every transmittance estimator within 4 sigma of exp(-tau) of the FP64 integrator;
  * white furnace (albedo 1): every path returns the constant sky;
  * the reference's relational equal-mean tests (test/TestVolumetricPathTracing.cpp:156-281) at 3 sigma;
  * supervoxel min/max tables against a NumPy brute force (SuperVoxelGrid.cpp:101-211, 269-405).
"""
import math

import numpy as np
import pytest


def test_hybrid_taus_seed0_is_bare_lcg(oracle):
    got = oracle.hybrid_taus(0, 6)
    w = 0
    seq = []
    for i in range(2 + 6):
        w = (1664525 * w + 1013904223) & 0xFFFFFFFF
        seq.append(w)
    expect = np.array([np.float32(np.float32(2.3283064365387e-10) * np.float32(v)) for v in seq[2:]], np.float32)
    np.testing.assert_array_equal(got, expect)


def test_hybrid_taus_uniform(oracle):
    x = np.concatenate([oracle.hybrid_taus(s, 64) for s in range(1, 4001)])
    assert abs(x.mean() - 0.5) < 4 * math.sqrt(1 / 12 / x.size)
    assert abs(x.var() - 1 / 12) < 0.002
    assert x.min() >= 0.0 and x.max() <= 1.0


def _ray_dir(ivp, sx, sy):
    def unproject(z):
        v = np.array([sx, sy, z, 1.0], np.float64)
        m = np.asarray(ivp, np.float64).reshape(4, 4).T  # column-major storage
        r = m @ v
        return r[:3] / r[3]

    p0, p1 = unproject(0.0), unproject(1.0)
    d = p1 - p0
    return p0, d / np.linalg.norm(d)


def test_camera_conventions(oracle):
    ivp, iv, _ = oracle.camera_matrices(aspect=1.0)
    p0, d = _ray_dir(ivp, 0.0, 0.0)
    np.testing.assert_allclose(d, [0, 0, -1], atol=1e-6)
    np.testing.assert_allclose(p0, [0, 0, 0.8 - 1.0 / 512.0], atol=1e-6)
    _, dtop = _ray_dir(ivp, 0.0, -1.0)  # NDC y = -1 is image row 0
    assert dtop[1] > 0, "row 0 must look up (top of the image)"
    assert abs(dtop[1] / -dtop[2] - 0.5) < 1e-5  # tan(fovy / 2) = 1/2
    _, dright = _ray_dir(ivp, 1.0, 0.0)
    assert dright[0] > 0 and abs(dright[0] / -dright[2] - 0.5) < 1e-5


def test_optical_depth_homogeneous_nearest(oracle):
    d = np.ones((8, 8, 8), np.float32)
    sc = oracle.Scene(d, interpolation="nearest", extinction=10.0)
    o = np.array([[-1.0, 0.01, 0.02], [0.05, -1.0, 0.0], [0.1, 0.1, 2.0]], np.float32)
    w = np.array([[1, 0, 0], [0, 1, 0], [0, 0, -1]], np.float32)
    tau = sc.optical_depth(o, w)
    np.testing.assert_allclose(tau, 10.0 * 0.5, rtol=1e-12)
    # diagonal chord of the [-0.25, 0.25]^3 cube
    o2 = np.array([[-1.0, -1.0, -1.0]], np.float32)
    w2 = np.array([[1, 1, 1]], np.float32) / np.float32(math.sqrt(3))
    np.testing.assert_allclose(sc.optical_depth(o2, w2), 10.0 * 0.5 * math.sqrt(3), rtol=1e-6)


def test_optical_depth_trilinear_border_ramp(oracle):
    n = 16
    d = np.ones((n, n, n), np.float32)
    sc = oracle.Scene(d, interpolation="trilinear", extinction=1.0)
    # through voxel centres along x at y, z in the interior: integral of the tent-filtered box = (n - 1/4) texels
    y = z = (5.5 / n) * 0.5 - 0.25
    tau = sc.optical_depth(np.array([[-1.0, y, z]], np.float32), np.array([[1.0, 0, 0]], np.float32))
    np.testing.assert_allclose(tau, (n - 0.25) * (0.5 / n), rtol=1e-6)


def _brute_tau(density, box, sigma, o, w, trilinear, steps=20000):
    nz, ny, nx = density.shape
    bmin, bmax = box
    t0, t1 = 0.0, 10.0
    for a in range(3):
        if abs(w[a]) < 1e-12:
            continue
        ta, tb = (bmin[a] - o[a]) / w[a], (bmax[a] - o[a]) / w[a]
        t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
    if t1 <= t0:
        return 0.0
    ts = t0 + (np.arange(steps) + 0.5) / steps * (t1 - t0)
    p = o[None, :] + ts[:, None] * w[None, :]
    u = (p - bmin) / (bmax - bmin) * np.array([nx, ny, nz])
    vmin, vmax = float(density.min()), float(density.max())
    pad = np.pad((density.astype(np.float64) - vmin) / (vmax - vmin), 1, constant_values=(0 - vmin) / (vmax - vmin))
    if not trilinear:
        i = np.floor(u).astype(int)
        val = pad[i[:, 2] + 1, i[:, 1] + 1, i[:, 0] + 1]
    else:
        uu = u - 0.5
        b = np.floor(uu).astype(int)
        f = uu - b
        val = 0.0
        for k in range(8):
            dx, dy, dz = k & 1, (k >> 1) & 1, (k >> 2) & 1
            wk = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * (f[:, 2] if dz else 1 - f[:, 2])
            val = val + wk * pad[b[:, 2] + dz + 1, b[:, 1] + dy + 1, b[:, 0] + dx + 1]
    return sigma * float(val.mean()) * (t1 - t0)


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_optical_depth_brute_force(oracle, interp):
    rng = np.random.default_rng(3)
    d = rng.random((6, 7, 5)).astype(np.float32)
    sc = oracle.Scene(d, interpolation=interp, extinction=3.0)
    box = (np.array(sc.sc.box_min[:], np.float64), np.array(sc.sc.box_max[:], np.float64))
    for _ in range(6):
        o = rng.uniform(-0.6, 0.6, 3)
        tgt = rng.uniform(-0.2, 0.2, 3)
        w = (tgt - o) / np.linalg.norm(tgt - o)
        tau = sc.optical_depth(o[None].astype(np.float32), w[None].astype(np.float32))[0]
        ref = _brute_tau(d, box, 3.0, o.astype(np.float32).astype(np.float64),
                         w.astype(np.float32).astype(np.float64), interp == "trilinear")
        assert abs(tau - ref) < 2e-3 * max(1.0, ref), (tau, ref)


def _rays_through_box(rng, n, box_half):
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * (1.5 * math.sqrt(3) * box_half)
    tgt = rng.uniform(-box_half, box_half, (n, 3))
    w = tgt - o
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    return o.astype(np.float32), w.astype(np.float32)


@pytest.mark.parametrize("estimator", ["delta", "ratio", "residual_ratio", "decomposition"])
@pytest.mark.parametrize("interp", ["nearest", "stochastic"])
def test_oracle_transmittance_unbiased(oracle, estimator, interp):
    from oracle.oracle import fbm_cloud

    d = fbm_cloud(16, 16, 16, seed=2)
    sc = oracle.Scene(d, interpolation=interp, extinction=24.0)
    rng = np.random.default_rng(11)
    o, w = _rays_through_box(rng, 48, 0.25)
    tau = sc.optical_depth(o, w)
    exact = np.exp(-tau)
    n = 4000
    mean, m2 = sc.transmittance(estimator, o, w, n_samples=n, seed=7)
    # sample variance, floored at the resolution 1/n of an n-sample mean of a binary estimator
    var = np.maximum(m2 - mean ** 2, 1.0 / n)
    z = (mean - exact) / np.sqrt(var / n)
    assert np.all(np.abs(z) < 5.0), z
    assert abs(z.mean()) * math.sqrt(z.size) < 4.0


def test_white_furnace(oracle):
    from oracle.oracle import fbm_cloud

    d = fbm_cloud(16, 16, 16, seed=1)
    sc = oracle.Scene(d, mode="delta", albedo=1.0, env="single_color", extinction=64.0, width=16, height=16)
    img, _ = sc.render(4)
    sky = np.array([2.6, 2.6 * 0.961538462, 2.6 * 0.884615385], np.float32)
    np.testing.assert_allclose(img, np.broadcast_to(sky, img.shape), rtol=1e-6)


def test_render_mean_sums_the_render_samples(oracle):
    """render_mean draws the samples render draws: one frame is the sample itself, the squared sums agree exactly,
    frame ranges add up, and the FP32 running mean stays within its rounding drift of the exact mean."""
    from oracle.oracle import fbm_cloud

    sc = oracle.Scene(fbm_cloud(12, 10, 8, seed=4), mode="delta", extinction=48.0, width=12, height=10)
    one, sq1 = sc.render(1)
    m1, q1 = sc.render_mean(1)
    np.testing.assert_array_equal(m1, one.astype(np.float64))
    np.testing.assert_array_equal(q1, sq1)
    acc, sq = sc.render(64)
    m, q = sc.render_mean(64)
    np.testing.assert_array_equal(q, sq)
    np.testing.assert_allclose(acc, m, rtol=1e-5, atol=1e-6)
    ma, _ = sc.render_mean(24)
    mb, _ = sc.render_mean(40, frame_start=24)
    np.testing.assert_allclose((24 * ma + 40 * mb) / 64, m, rtol=1e-12, atol=1e-15)


def _mean_and_sigma(img, sq, frames):
    """Image mean per channel and its standard error from the per-pixel sample variances."""
    npix = img.shape[0] * img.shape[1]
    var = np.maximum(sq / frames - img.astype(np.float64) ** 2, 0.0)
    return img.reshape(-1, 3).mean(0), np.sqrt(var.reshape(-1, 3).sum(0) / frames) / npix


def _equal_mean(a, b, frames):
    ma, sa = _mean_and_sigma(*a, frames)
    mb, sb = _mean_and_sigma(*b, frames)
    tol = 3.0 * np.sqrt(sa ** 2 + sb ** 2) + 1e-6
    assert np.all(np.abs(ma - mb) <= tol), (ma, mb, tol)


def _block(n, boundary=False):
    d = np.ones((n, n, n), np.float32)
    if boundary and n >= 3:
        d[:] = 0
        d[1:-1, 1:-1, 1:-1] = 1
    return d


RES, SPP = 24, 32


def test_relational_delta_vs_ratio(oracle):
    # TestVolumetricPathTracing.cpp:156-164 (1^3 block of density 1)
    a = oracle.Scene(_block(1), mode="delta", width=RES, height=RES, extinction=64.0).render(SPP)
    b = oracle.Scene(_block(1), mode="ratio", width=RES, height=RES, extinction=64.0).render(SPP)
    _equal_mean(a, b, SPP)


def test_relational_seed_offset(oracle):
    # TestVolumetricPathTracing.cpp:165-174
    a = oracle.Scene(_block(1), mode="delta", width=RES, height=RES, extinction=64.0).render(SPP)
    b = oracle.Scene(_block(1), mode="delta", width=RES, height=RES, extinction=64.0,
                     seed_offset=268435456).render(SPP)
    _equal_mean(a, b, SPP)


def test_relational_delta_vs_decomposition(oracle):
    # TestVolumetricPathTracing.cpp:239-271 (8^3 block, nearest, with an empty boundary layer)
    a = oracle.Scene(_block(8, True), mode="delta", interpolation="nearest", width=RES, height=RES,
                     extinction=64.0).render(SPP)
    b = oracle.Scene(_block(8, True), mode="decomposition", interpolation="nearest", width=RES, height=RES,
                     extinction=64.0).render(SPP)
    _equal_mean(a, b, SPP)


def test_relational_delta_vs_next_event(oracle):
    # TestVolumetricPathTracing.cpp:273-281
    a = oracle.Scene(_block(8, True), mode="delta", width=RES, height=RES, extinction=64.0).render(SPP)
    b = oracle.Scene(_block(8, True), mode="next_event", width=RES, height=RES, extinction=64.0).render(SPP)
    _equal_mean(a, b, SPP)


def test_next_event_env_plus_light_split(oracle):
    """NEE with an environment image and a directional light (50/50 choice, factor 2, NextEventTracking.glsl:
    318-345) equals, in expectation, the environment-only render plus the light-only render (black environment:
    the light is always sampled with factor 1)."""
    d = _block(8, True)
    env = np.ones((3, 8, 16), np.float32)
    L = oracle.make_light(light_type=2, light_space=0, position=(0.0, 100.0, 0.0), intensity=2.0)
    both = oracle.Scene(d, mode="next_event", env_image=env, lights=[L], width=RES, height=RES,
                        extinction=64.0).render(SPP)
    env_only = oracle.Scene(d, mode="next_event", env_image=env, width=RES, height=RES,
                            extinction=64.0).render(SPP)
    light_only = oracle.Scene(d, mode="next_event", env_image=env, env_factor=(0.0, 0.0, 0.0), lights=[L],
                              width=RES, height=RES, extinction=64.0).render(SPP)
    # E[both] = E[env part] + E[light part]; the light-only render still adds the (zero) env for escaping rays.
    mb, sb = _mean_and_sigma(*both, SPP)
    me, se = _mean_and_sigma(*env_only, SPP)
    ml, sl = _mean_and_sigma(*light_only, SPP)
    tol = 3.0 * np.sqrt(sb ** 2 + se ** 2 + sl ** 2) + 1e-6
    assert np.all(np.abs(mb - (me + ml)) <= tol), (mb, me + ml, tol)


@pytest.mark.parametrize("interp", ["nearest", "stochastic"])
def test_supervoxel_tables_min_max(oracle, interp):
    rng = np.random.default_rng(5)
    d = (rng.random((20, 12, 16)) * (rng.random((20, 12, 16)) > 0.5)).astype(np.float32)
    sc = oracle.Scene(d, interpolation=interp, extinction=100.0)
    sv, g, rrt, dec, occ = sc.supervoxel_tables()
    assert sv == 8 and g == (2, 2, 3)
    vmin, vmax = float(d.min()), float(d.max())
    norm = (d - vmin) / (vmax - vmin)
    halo = 0 if interp == "nearest" else 1
    pad = np.pad(norm, halo, constant_values=0.0) if halo else norm
    for si in range(g[0] * g[1] * g[2]):
        x, y, z = si % g[0], (si // g[0]) % g[1], si // (g[0] * g[1])
        lo = [x * 8, y * 8, z * 8]
        block = np.zeros((8 + 2 * halo,) * 3, np.float32)
        # zero outside the grid, like the reference's clamp-to-zero border
        src = pad[lo[2]:lo[2] + 8 + 2 * halo, lo[1]:lo[1] + 8 + 2 * halo, lo[0]:lo[0] + 8 + 2 * halo]
        block[:src.shape[0], :src.shape[1], :src.shape[2]] = src
        assert dec[si, 0] == block.min() and dec[si, 1] == block.max()
        assert occ[si] == (1 if block.max() >= 1e-5 else 0)
        partial = ((x == g[0] - 1 and d.shape[2] % 8) or (y == g[1] - 1 and d.shape[1] % 8)
                   or (z == g[2] - 1 and d.shape[0] % 8))
        mu_min = 0.0 if partial else np.float32(block.min()) * np.float32(100.0)
        mu_r = max(np.float32(block.max()) * np.float32(100.0) - np.float32(mu_min), np.float32(0.1))
        assert rrt[si, 1] == np.float32(mu_r)
        assert mu_min - 1e-4 <= rrt[si, 0] <= 100.0 * block.mean() + 1e-3


@pytest.mark.parametrize("mode", ["delta", "ratio", "next_event", "decomposition"])
def test_max_bounces_truncates_paths(oracle, mode):
    """The bounce limit (a workload bound the reference does not have) only cuts paths: with the same random stream a
    limit far above any path's length renders bit for bit what the unbounded tracker renders, and for the estimators
    whose radiance is only collected at the path's end (delta, ratio, decomposition) a cut path gives 0, so every
    pixel of a limited render is at most the unbounded one and the limits are ordered."""
    d = oracle.fbm_cloud(16, 16, 16, seed=21)
    kw = dict(mode=mode, extinction=96.0, g=0.5, width=16, height=16)
    free, _ = oracle.Scene(d, **kw).render(8)
    high, _ = oracle.Scene(d, max_bounces=100000, **kw).render(8)
    np.testing.assert_array_equal(free, high)
    one, _ = oracle.Scene(d, max_bounces=1, **kw).render(8)
    two, _ = oracle.Scene(d, max_bounces=2, **kw).render(8)
    assert not np.array_equal(one, free)
    if mode != "next_event":
        # running means in FP32: allow the last-ulp difference of the 1/(frame+1) mix
        assert np.all(one <= two * (1 + 1e-6) + 1e-7)
        assert np.all(two <= free * (1 + 1e-6) + 1e-7)


@pytest.mark.parametrize("spectral,scalar", [("delta_spectral", "delta"), ("next_event_spectral", "next_event")])
def test_spectral_with_grey_medium_is_the_scalar_tracker(oracle, spectral, scalar):
    """With grey extinction and albedo every spectral weight update is sigma / (majorant * P) = 1 and the collision
    probabilities are the scalar tracker's, so both consume the same random numbers and agree pixel for pixel up to
    FP32 rounding of the normalised probabilities."""
    d = oracle.fbm_cloud(16, 16, 16, seed=23)
    kw = dict(extinction=80.0, albedo=0.8, g=0.5, width=24, height=24)
    a, _ = oracle.Scene(d, mode=spectral, **kw).render(16)
    b, _ = oracle.Scene(d, mode=scalar, **kw).render(16)
    close = np.abs(a - b) <= 1e-4 * np.maximum(1.0, np.abs(b))
    assert close.mean() > 0.99, close.mean()
    np.testing.assert_allclose(a.mean(), b.mean(), rtol=2e-3)


def test_spectral_white_furnace_per_channel(oracle):
    """Albedo 1 with a constant sky: each channel of spectral tracking is an unbiased estimate of the sky radiance,
    whatever the chromatic extinction (Kutz et al. 2017)."""
    d = oracle.fbm_cloud(16, 16, 16, seed=24)
    sc = oracle.Scene(d, mode="delta_spectral", extinction=(120.0, 80.0, 40.0), albedo=1.0, g=0.3, env="single_color",
                      width=32, height=32)
    img, sq = sc.render(64)
    n = 32 * 32 * 64
    mean = img.reshape(-1, 3).mean(0).astype(np.float64)
    var = np.maximum(sq.reshape(-1, 3).sum(0) / n - mean ** 2, 1e-12)
    sky = np.array([2.6, 2.6 * 0.961538462, 2.6 * 0.884615385])
    z = (mean - sky) / np.sqrt(var / n)
    assert np.all(np.abs(z) < 4.0), (mean, sky, z)


def _feature_scene(oracle, density, **kw):
    args = dict(mode="delta", interpolation="nearest", extinction=2.0e4, albedo=0.9, g=0.5, width=24, height=24)
    args.update(kw)
    return oracle.Scene(density, **args)


def test_feature_maps_share_one_first_event(oracle):
    """Clouds.glsl writes every map from the same first event: the coverage channels of first X, albedo and cloud
    only agree, albedo = scattering albedo x coverage, and depth far = depth + (1 - coverage) x far distance."""
    d = oracle.fbm_cloud(20, 20, 20, seed=31)
    sc = _feature_scene(oracle, d, extinction=60.0)
    m = sc.render_features(12)
    cov = m["first_x"][..., 3]
    assert 0.05 < cov.mean() < 0.95
    np.testing.assert_allclose(m["albedo"][..., 3], cov, atol=1e-6)
    np.testing.assert_allclose(m["cloud_only"][..., 3], cov, atol=1e-6)
    np.testing.assert_allclose(m["albedo"][..., 0], 0.9 * cov, atol=1e-6)
    np.testing.assert_allclose(m["depth_far"][..., 0], m["depth"][..., 0] + (1.0 - cov) * 80.0, rtol=1e-5, atol=1e-4)
    # results equal the plain render's running mean
    img, _ = sc.render(12)
    np.testing.assert_allclose(m["result"][..., :3], img, rtol=1e-6, atol=1e-7)


def test_feature_maps_opaque_medium(oracle):
    """A dense absorbing cube stops every ray at the face it enters (z = 0.25 for the camera at z = 0.8 looking down
    -z): first X lies on that face, depth is the distance from the ray's near-plane origin, the normal is the
    central difference of getCloudFiniteDifference across the face, (0, 0, 1) away from the cube's side edges (the
    border outside the grid is 0), and with the previous view-projection equal to the current one the flow of a
    pixel is its sub-pixel jitter offset (|flow| <= 0.5) and the reprojected UV lies in the pixel's footprint."""
    n = 16
    cube = np.ones((n, n, n), np.float32)
    cube[0, 0, 0] = 0.0  # a non-constant field (normalisation range [0, 1]) that leaves the +z face untouched
    sc = _feature_scene(oracle, cube, albedo=0.0, width=16, height=16)
    m = sc.render_features(8)
    hit = m["first_x"][..., 3] == 1.0
    assert hit.sum() > 20
    fx = m["first_x"][hit][:, :3]
    np.testing.assert_allclose(fx[:, 2], 0.25, atol=2e-3)
    r = np.linalg.norm(fx - np.array([0.0, 0.0, 0.8]), axis=1)
    cos = (0.8 - fx[:, 2]) / r
    np.testing.assert_allclose(m["depth"][..., 0][hit], r - (1.0 / 512.0) / cos, atol=2e-3)
    inner = hit & (np.abs(m["first_x"][..., 0]) < 0.2) & (np.abs(m["first_x"][..., 1]) < 0.2)
    assert inner.sum() > 10
    nrm = m["normal"][inner][:, :3]
    np.testing.assert_allclose(nrm, np.broadcast_to([0.0, 0.0, 1.0], nrm.shape), atol=1e-5)
    assert np.all(np.abs(m["flow"][hit]) <= 0.5 + 1e-4)
    ys, xs = np.nonzero(hit)
    u = m["reproj_uv"][ys, xs, 0] * 16
    assert np.all(u >= xs - 1e-3) and np.all(u <= xs + 1 + 1e-3)
