"""Known-answer pins of the oracle functions that the statement-by-statement restatement alone cannot vouch for:
the residual-ratio radiance loop (ResidualRatioTracking.glsl:78-224), the DEFAULT sky and its sun lobe
(VptUtils.glsl:335-391), point / spot / directional lights (VptUtils.glsl:394-453) and the residual-ratio control
extinction mu_c (SuperVoxelGrid.cpp:218-256).

Each expected value comes from a closed form or from the reference's literal constants, never from the oracle's
own code path, and each check is chosen so that a plausible slip (a dropped factor, a wrong sign, a swapped stop,
a wrong exponent or continuation rule) fails it.

Statistical checks use a family-wise false-alarm rate of 0.27 % (the two-sided 3-sigma level) split over the k
comparisons of the test (Bonferroni): each |z| must stay below z_k = Phi^-1(1 - 0.0027 / (2k)), 3.0 for k = 1.
"""
import math

import numpy as np
import pytest
from scipy.stats import norm

ALPHA = 0.0027


def z_bound(k):
    return float(norm.isf(ALPHA / (2.0 * k)))


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _rays_through_box(rng, n, inner=0.2, radius=1.0):
    """Origins on a sphere of the given radius, directions towards random points of the inner box [-inner, inner]^3."""
    o = _unit(rng.normal(size=(n, 3))) * radius
    t = rng.uniform(-inner, inner, size=(n, 3))
    return o.astype(np.float32), _unit(t - o).astype(np.float32)


def _slab(o, d, bmin, bmax):
    """FP64 slab test: (t_entry, t_exit) per ray (t_entry clamped to 0)."""
    o = o.astype(np.float64)
    d = d.astype(np.float64)
    with np.errstate(divide="ignore"):
        inv = 1.0 / d
    t0 = (np.asarray(bmin, np.float64) - o) * inv
    t1 = (np.asarray(bmax, np.float64) - o) * inv
    tn = np.maximum(np.minimum(t0, t1).max(1), 0.0)
    tf = np.maximum(t0, t1).min(1)
    return tn, tf


SKY = (1.0, 0.5, 0.25)  # Single Color sky = sun intensity 1 x colour (1, 0.5, 0.25): sums of it stay exact


def _rrt_scene(oracle, density, interp="nearest", ext=8.0, albedo=0.9, voxel_range=None, g=0.5):
    return oracle.Scene(density, mode="residual_ratio", interpolation=interp, extinction=ext, albedo=albedo, g=g,
                        env="single_color", sun_intensity=1.0, sun_color=SKY, super_voxel_size=8,
                        voxel_range=voxel_range)


# ---------------------------------------------------------------------------------------------------------------
# Residual ratio tracking radiance loop
# ---------------------------------------------------------------------------------------------------------------

def test_rrt_first_traversal_homogeneous_closed_form(oracle):
    """Homogeneous cube, nearest filtering, grid a multiple of the supervoxel edge: every supervoxel has
    mu_min = mu_max = mu_avg, so mu_c = mu and each residual factor 1 - (mu - mu_c) / mu_r_bar is exactly 1. The
    first traversal alone (no in-scattering iteration) then returns exp(-mu L) x sky deterministically, with L the
    chord from the entry into the box shrunk by 1e-6 to the exit from the box."""
    rng = np.random.default_rng(11)
    d = np.ones((16, 16, 16), np.float32)
    ext = 8.0
    sc = _rrt_scene(oracle, d, ext=ext)
    o, w = _rays_through_box(rng, 4000)
    rgb, it = sc.trace_rays(o, w, seed=5, rrt_max_iterations=0)
    assert np.all(it == 0)
    tn, _ = _slab(o, w, [-0.25 + 1e-6] * 3, [0.25 - 1e-6] * 3)
    x = o.astype(np.float64) + w.astype(np.float64) * tn[:, None]
    _, tf = _slab(x, w, [-0.25] * 3, [0.25] * 3)
    expect = np.exp(-ext * tf)[:, None] * np.array(SKY)
    np.testing.assert_allclose(rgb, expect, rtol=2e-5, atol=0.0)


def test_rrt_first_traversal_unbiased_vs_fp64(oracle):
    """fBm cloud with stochastic filtering: the first traversal's sky term is T x sky, and T is the residual ratio
    estimate of the transmittance, so its mean over samples equals exp(-tau) of the FP64 optical depth per ray."""
    rng = np.random.default_rng(12)
    dens = oracle.fbm_cloud(24, 24, 24, seed=9)
    sc = _rrt_scene(oracle, dens, interp="stochastic", ext=48.0)
    n_rays, n_samp = 120, 600
    o, w = _rays_through_box(rng, n_rays)
    tau = sc.optical_depth(o, w)
    rgb, _ = sc.trace_rays(np.repeat(o, n_samp, 0), np.repeat(w, n_samp, 0), seed=77, rrt_max_iterations=0)
    t = rgb[:, 0].astype(np.float64).reshape(n_rays, n_samp) / SKY[0]
    np.testing.assert_allclose(rgb[:, 1] / SKY[1], rgb[:, 0] / SKY[0], rtol=1e-6)
    m = t.mean(1)
    se = t.std(1, ddof=1) / math.sqrt(n_samp)
    expect = np.exp(-tau)
    z = (m - expect) / np.maximum(se, 1e-6)
    assert np.all(np.abs(z) < z_bound(n_rays + 1)), float(np.abs(z).max())
    z_all = (m - expect).sum() / math.sqrt((se ** 2).sum())
    assert abs(z_all) < z_bound(n_rays + 1), z_all


def test_rrt_empty_space_in_scattering_closed_form(oracle):
    """Zero density (normalisation range [0, 1]): every supervoxel has mu_c = 0 and the floor mu_r_bar = 0.1, so
    tentative collisions arrive at rate 0.1 per unit length with weight exactly 1, and the path's radiance is
    exactly sky x (1 + K) for K in-scattering iterations. A traversal continues iff it saw a tentative collision
    (the reservoir weight sum is then >= 1 > xi), so the first one continues with probability
    1 - exp(-0.1 L) for its chord L."""
    rng = np.random.default_rng(13)
    d = np.zeros((16, 16, 16), np.float32)
    sc = _rrt_scene(oracle, d, voxel_range=(0.0, 1.0), ext=1024.0)
    n = 40000
    o, w = _rays_through_box(rng, n)
    rgb, it = sc.trace_rays(o, w, seed=3)
    np.testing.assert_array_equal(rgb, (1.0 + it.astype(np.float32))[:, None] * np.array(SKY, np.float32))
    assert it.max() <= 10
    tn, _ = _slab(o, w, [-0.25 + 1e-6] * 3, [0.25 - 1e-6] * 3)
    x = o.astype(np.float64) + w.astype(np.float64) * tn[:, None]
    _, tf = _slab(x, w, [-0.25] * 3, [0.25] * 3)
    p = 1.0 - np.exp(-0.1 * tf)
    k = int((it >= 1).sum())
    z = (k - p.sum()) / math.sqrt((p * (1.0 - p)).sum())
    assert abs(z) < z_bound(1), (k, p.sum(), z)


def test_rrt_ignores_albedo_bit_exact(oracle):
    """The reference's residual ratio tracker never reads the scattering albedo (ResidualRatioTracking.glsl:86-224
    computes absorptionAlbedo and Ps but uses neither), so images with different albedos are identical."""
    dens = oracle.fbm_cloud(20, 16, 12, seed=2)
    imgs = []
    for a in (0.9, 0.3):
        sc = oracle.Scene(dens, mode="residual_ratio", interpolation="stochastic", extinction=64.0, albedo=a,
                          g=0.5, width=16, height=16)
        img, _ = sc.render(8)
        imgs.append(img)
    np.testing.assert_array_equal(imgs[0], imgs[1])


# ---------------------------------------------------------------------------------------------------------------
# DEFAULT sky and sun lobe (VptUtils.glsl:335-391)
# ---------------------------------------------------------------------------------------------------------------

# The reference's gradient stops: heights dir.y and colours (VptUtils.glsl:339-351).
STOPS = [(-1.0, (0.1, 0.05, 0.01)), (-0.1, (0.01, 0.05, 0.2)), (0.0, (0.8, 0.9, 1.0)), (0.4, (0.1, 0.3, 1.0)),
         (1.0, (0.01, 0.1, 0.7))]


def _srgb_to_linear(c):
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


@pytest.mark.parametrize("linear", [False, True])
def test_default_sky_gradient_stops(oracle, linear):
    """At each stop height the gradient equals the stop colour (sun far away: its lobe is zero); linear RGB output
    applies the sRGB transfer function (IEC 61966-2-1) to it."""
    sc = oracle.Scene(np.ones((4, 4, 4), np.float32), env="default", sun_direction=(1.0, 0.0, 0.0),
                      use_linear_rgb=linear)
    dirs = np.array([[0.0, y, -math.sqrt(max(0.0, 1.0 - y * y))] for y, _ in STOPS], np.float32)
    got = sc.sky(dirs)
    expect = np.array([c for _, c in STOPS], np.float64)
    if linear:
        expect = _srgb_to_linear(expect)
    np.testing.assert_allclose(got, expect, rtol=2e-6, atol=1e-7)
    # halfway between the stops at 0 and 0.4 in height: smoothstep(0.5) = 0.5 mixes the two colours equally
    mid = sc.sky(np.array([[0.0, 0.2, -math.sqrt(1.0 - 0.04)]], np.float32))[0]
    m = 0.5 * (np.array(STOPS[2][1]) + np.array(STOPS[3][1]))
    np.testing.assert_allclose(mid, _srgb_to_linear(m) if linear else m, rtol=2e-6)


def test_sun_lobe_phong(oracle):
    """sampleLight: sun intensity x max(0, cos)^10 x (10 + 1) / (2 x 3.14159) (the reference's Phong lobe with its
    own truncated pi), zero behind the sun."""
    sun_dir = _unit([0.5826, 0.7660, 0.2717])
    sc = oracle.Scene(np.ones((4, 4, 4), np.float32), env="default", sun_direction=tuple(sun_dir),
                      use_linear_rgb=False, sun_intensity=2.6)
    sun = 2.6 * np.array([1.0, 0.961538462, 0.884615385])
    t = _unit(np.cross(sun_dir, [0.0, 0.0, 1.0]))
    dirs = []
    cosines = [1.0, 0.5, 0.0, -0.5]
    for c in cosines:
        dirs.append(c * sun_dir + math.sqrt(1.0 - c * c) * t)
    dirs = np.array(dirs, np.float32)
    got = sc.sky(dirs)
    # the gradient part alone: the same directions with the sun moved opposite (lobe zero for cos >= 0 dirs)
    far = oracle.Scene(np.ones((4, 4, 4), np.float32), env="default", sun_direction=tuple(-sun_dir),
                       use_linear_rgb=False, sun_intensity=2.6)
    grad = far.sky(dirs[:3])
    lobe = got[:3] - grad
    expect = np.array([max(0.0, c) ** 10 * 11.0 / (2.0 * 3.14159) for c in cosines[:3]])[:, None] * sun
    np.testing.assert_allclose(lobe, expect, rtol=2e-5, atol=2e-6)
    # behind the sun (cos < 0) the lobe is zero: the sky equals the mirrored-sun scene's gradient there
    np.testing.assert_allclose(got[3], oracle.Scene(np.ones((4, 4, 4), np.float32), env="default",
                                                    sun_direction=tuple(sun_dir), use_linear_rgb=False,
                                                    sun_intensity=0.0).sky(dirs[3:4])[0], rtol=1e-6)


# ---------------------------------------------------------------------------------------------------------------
# Lights (VptUtils.glsl:394-453)
# ---------------------------------------------------------------------------------------------------------------

def _light_scene(oracle, light):
    return oracle.Scene(np.ones((4, 4, 4), np.float32), env="black", lights=[light])


def test_point_light_inverse_square(oracle):
    p = np.array([0.1, 0.2, 0.3])
    L = oracle.make_light(light_type=0, light_space=0, color=(1.0, 0.5, 0.25), intensity=3.0, position=tuple(p))
    sc = _light_scene(oracle, L)
    dirs = _unit(np.random.default_rng(4).normal(size=(64, 3)))
    for dist in (0.5, 2.0):
        x = p + dirs * dist
        got = sc.light_radiance(0, x)
        np.testing.assert_allclose(got, np.broadcast_to(3.0 / dist ** 2 * np.array([1.0, 0.5, 0.25]), got.shape),
                                   rtol=2e-6)
    # the distance falloff is clamped at 1e-3 squared units; without use_distance there is none
    got = sc.light_radiance(0, (p + np.array([0.01, 0.0, 0.0]))[None])[0]
    np.testing.assert_allclose(got, 3.0 / 1e-3 * np.array([1.0, 0.5, 0.25]), rtol=2e-6)
    L2 = oracle.make_light(light_type=0, light_space=0, color=(1.0, 0.5, 0.25), intensity=3.0, position=tuple(p),
                           use_distance=0)
    got = _light_scene(oracle, L2).light_radiance(0, (p + 0.7 * dirs)[:8])
    np.testing.assert_allclose(got, np.broadcast_to(3.0 * np.array([1.0, 0.5, 0.25]), got.shape), rtol=1e-7)


def test_spot_light_cone(oracle):
    """Full intensity inside the falloff cone, zero outside the total width, smoothstep(cos total, cos falloff,
    cos theta) between: 0.5 where cos theta is halfway between the two cosines."""
    p = np.array([0.0, 0.0, 0.5])
    tot, fall = 0.4, 0.2
    L = oracle.make_light(light_type=1, light_space=0, color=(1.0, 1.0, 1.0), intensity=2.0, position=tuple(p),
                          spot_total_width=tot, spot_falloff_start=fall, spot_direction=(0.0, 0.0, -1.0))
    sc = _light_scene(oracle, L)
    c_mid = 0.5 * (math.cos(tot) + math.cos(fall))
    cases = [(0.0, 1.0), (0.1, 1.0), (0.5, 0.0), (math.acos(c_mid), 0.5)]
    for theta, frac in cases:
        x = p + 1.0 * np.array([math.sin(theta), 0.0, -math.cos(theta)])
        got = sc.light_radiance(0, x[None])[0]
        np.testing.assert_allclose(got, [2.0 * frac] * 3, rtol=1e-4, atol=1e-6)


def test_directional_light_constant(oracle):
    L = oracle.make_light(light_type=2, light_space=0, color=(0.2, 0.4, 0.8), intensity=2.5, position=(0.0, 1.0, 0.0))
    got = _light_scene(oracle, L).light_radiance(0, np.random.default_rng(6).normal(size=(16, 3)))
    np.testing.assert_allclose(got, np.broadcast_to(2.5 * np.array([0.2, 0.4, 0.8]), got.shape), rtol=1e-7)


# ---------------------------------------------------------------------------------------------------------------
# Control extinction of residual ratio tracking (SuperVoxelGrid.cpp:218-256)
# ---------------------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("ext", [1.0, 64.0, 1024.0])
def test_mu_c_closed_form(oracle, ext):
    """Novak et al. 2014's control coefficient as the reference sets it: mu_r_bar = max(mu_max - mu_min, 0.1),
    mu_c = clamp(mu_min + mu_r_bar * gamma^(1 / (D mu_r_bar) - 1), mu_min, mu_avg) with gamma = 2 and D = sqrt(3) S.
    A supervoxel three quarters full (mu_avg = 0.75 ext) leaves the clamp inactive, so the exponent is visible; a
    homogeneous one gives mu_c = mu."""
    d = np.zeros((16, 16, 16), np.float32)
    d[0:8, 0:8, 0:8] = 1.0
    d[0:2, 0:8, 0:8] = 0.0  # supervoxel (0, 0, 0): 3/4 ones, 1/4 zeros
    d[8:16, 8:16, 8:16] = 1.0  # supervoxel (1, 1, 1): homogeneous
    sc = oracle.Scene(d, mode="residual_ratio", interpolation="nearest", extinction=ext, super_voxel_size=8)
    sv, grid, rrt, dec, _ = sc.supervoxel_tables()
    assert sv == 8 and grid == (2, 2, 2)
    D = math.sqrt(3.0) * 8.0
    mu_r = ext
    mu_c = min(max(mu_r * 2.0 ** (1.0 / (D * mu_r) - 1.0), 0.0), 0.75 * ext)
    assert mu_c < 0.75 * ext
    np.testing.assert_allclose(rrt[0], [mu_c, mu_r], rtol=3e-7)
    np.testing.assert_array_equal(rrt[7], np.float32([ext, 0.1]))
    np.testing.assert_array_equal(dec[0], [0.0, 1.0])
    np.testing.assert_array_equal(dec[7], [1.0, 1.0])
    # an empty supervoxel: mu_c = 0 and the floor mu_r_bar = 0.1
    np.testing.assert_array_equal(rrt[1], np.float32([0.0, 0.1]))


def test_zcrit_family_correction():
    """The tests' critical values: one comparison at 3 sigma, Bonferroni for families."""
    from _stats import zcrit

    assert abs(zcrit(1) - 3.0) < 1e-3
    assert abs(zcrit(3) - 3.32) < 0.01
    assert zcrit(512) > zcrit(12) > zcrit(3)
