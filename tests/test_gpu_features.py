"""GPU checks of the feature maps (Clouds.glsl:146-383) against the oracle's restatement: exact relations between the
maps of one first event, an opaque cube whose answers are geometric, and image means of every map against the
oracle's (24 seed-offset batches on each side; Student's t at the Bonferroni-corrected 3-sigma level)."""
import math

import numpy as np
import pytest

from _stats import tcrit

pytestmark = pytest.mark.gpu

RES = 24
NAMES = ["Result", "First X", "First W", "Normal", "Cloud Only", "Depth", "Depth Far", "Density", "Background",
         "Reprojected UV", "Albedo", "Flow", "Flow Reverse", "Depth (nabla)", "Depth (fwidth)"]
ORACLE = {"First X": "first_x", "Normal": "normal", "Cloud Only": "cloud_only", "Depth": "depth",
          "Depth Far": "depth_far", "Density": "density", "Background": "background", "Albedo": "albedo",
          "Reprojected UV": "reproj_uv", "Flow Reverse": "flow_reverse"}
# "Depth (nabla)" / "Depth (fwidth)" divide by sqrt(1 - A^2) of the view-space normal and are infinite wherever a
# normal lies along the view x or y axis, on the GPU as in the reference; they are checked on the opaque cube.


def _torch():
    import torch

    return torch


def _maps(vpt, frames, names=NAMES, res=RES):
    torch = _torch()
    t = torch.zeros((4, res, res), dtype=torch.float32, device="cuda")
    vpt.render_frame(t, frames)
    return {n: vpt.get_feature_map_from_string(t, n).cpu().numpy() for n in names}


def _setup(vpt, d, mode, interp="stochastic", ext=60.0, albedo=0.9):
    vpt.set_density_field(d)
    vpt.set_vpt_mode(mode)
    vpt.set_grid_interpolation_type(interp)
    vpt.set_extinction_scale(ext)
    vpt.set_extinction_base([1.0, 1.0, 1.0])
    vpt.set_scattering_albedo([albedo] * 3)
    vpt.set_phase_g(0.5)
    vpt.set_use_builtin_environment_map("Default")
    vpt.set_use_lights(False)
    vpt.set_seed_offset(0)
    vpt.set_use_feature_maps(NAMES)


def test_maps_share_one_first_event(vpt, oracle):
    print("case: feature map relations delta 20^3", flush=True)
    _setup(vpt, oracle.fbm_cloud(20, 20, 20, seed=31), 0)
    m = _maps(vpt, 12)
    cov = m["First X"][3]
    assert 0.05 < cov.mean() < 0.95
    np.testing.assert_allclose(m["Albedo"][3], cov, atol=1e-6)
    np.testing.assert_allclose(m["Cloud Only"][3], cov, atol=1e-6)
    np.testing.assert_allclose(m["Albedo"][0], 0.9 * cov, atol=1e-6)
    np.testing.assert_allclose(m["Depth Far"][0], m["Depth"][0] + (1.0 - cov) * 80.0, rtol=1e-5, atol=1e-3)
    torch = _torch()
    vpt.set_seed_offset(0)
    img = vpt.render_frame(torch.zeros((3, RES, RES), dtype=torch.float32, device="cuda"), 12).cpu().numpy()
    res = vpt.get_feature_map_from_string(torch.zeros((3, RES, RES), device="cuda"), "Result").cpu().numpy()
    np.testing.assert_array_equal(img, res)
    vpt.set_feature_map_type(6)  # render_frame returns "Depth"
    vpt.set_seed_offset(0)
    dep = vpt.render_frame(torch.zeros((4, RES, RES), dtype=torch.float32, device="cuda"), 12).cpu().numpy()
    vpt.set_feature_map_type(0)
    np.testing.assert_allclose(dep[:2], m["Depth"][:2], rtol=1e-6, atol=1e-6)
    assert np.all(dep[2] == 0.0) and np.all(dep[3] == 1.0)


def test_opaque_cube(vpt, oracle):
    print("case: feature maps opaque cube 16^3", flush=True)
    n = 16
    cube = np.ones((n, n, n), np.float32)
    cube[0, 0, 0] = 0.0
    _setup(vpt, cube, 0, interp="nearest", ext=2.0e4, albedo=0.0)
    m = _maps(vpt, 8, res=16)
    hit = m["First X"][3] == 1.0
    assert hit.sum() > 20
    fx = np.stack([m["First X"][i][hit] for i in range(3)], 1)
    np.testing.assert_allclose(fx[:, 2], 0.25, atol=2e-3)
    r = np.linalg.norm(fx - np.array([0.0, 0.0, 0.8]), axis=1)
    cos = (0.8 - fx[:, 2]) / r
    np.testing.assert_allclose(m["Depth"][0][hit], r - (1.0 / 512.0) / cos, atol=2e-3)
    inner = hit & (np.abs(m["First X"][0]) < 0.2) & (np.abs(m["First X"][1]) < 0.2)
    assert inner.sum() > 10
    nrm = np.stack([m["Normal"][i][inner] for i in range(3)], 1)
    np.testing.assert_allclose(nrm, np.broadcast_to([0.0, 0.0, 1.0], nrm.shape), atol=1e-5)
    assert np.all(np.abs(m["Flow"][:2][:, hit]) <= 0.5 + 1e-4)
    # the +z face normal is the view axis: A = B = 0, so nabla and fwidth vanish
    np.testing.assert_allclose(m["Depth (nabla)"][:2][:, inner], 0.0, atol=1e-5)
    np.testing.assert_allclose(m["Depth (fwidth)"][0][inner], 0.0, atol=1e-5)
    ys, xs = np.nonzero(hit)
    u = m["Reprojected UV"][0][ys, xs] * 16
    assert np.all(u >= xs - 1e-3) and np.all(u <= xs + 1 + 1e-3)


@pytest.mark.parametrize("name,mode", [("delta", 0), ("next_event", 5)])
def test_map_means_vs_oracle(vpt, oracle, name, mode):
    print(f"case: feature map means vs oracle {name}", flush=True)
    d = oracle.fbm_cloud(24, 20, 16, seed=33)
    _setup(vpt, d, mode)
    gpu = {k: [] for k in ORACLE}
    nb = 24
    for b in range(nb):
        vpt.set_seed_offset(1000003 * (b + 1))
        m = _maps(vpt, 32, names=list(ORACLE))
        for k in ORACLE:
            gpu[k].append(m[k].reshape(m[k].shape[0], -1).mean(1))
    vpt.set_seed_offset(0)
    ref = {k: [] for k in ORACLE}
    for b in range(nb):
        sc = oracle.Scene(d, mode=name, extinction=60.0, g=0.5, width=RES, height=RES, seed_offset=7919 * (b + 1))
        o = sc.render_features(32, names=list(ORACLE.values()))
        for k, on in ORACLE.items():
            ref[k].append(o[on].reshape(-1, o[on].shape[-1]).mean(0))
    # one family over every channel of every map; batch standard errors -> Student's t with Welch's dof
    comps = sum(np.array(ref[k]).shape[1] for k in ORACLE)
    for k in ORACLE:
        g = np.array(gpu[k], np.float64)
        r = np.array(ref[k], np.float64)
        c = r.shape[1]
        g = g[:, :c]
        mg, sg = g.mean(0), g.std(0, ddof=1) / math.sqrt(nb)
        mr, sr = r.mean(0), r.std(0, ddof=1) / math.sqrt(nb)
        se2 = sg ** 2 + sr ** 2
        dof = se2 ** 2 / np.maximum(sg ** 4 / (nb - 1) + sr ** 4 / (nb - 1), 1e-300)
        crit = np.array([tcrit(comps, max(float(v), 1.0)) for v in dof])
        z = (mg - mr) / np.maximum(np.sqrt(se2), 1e-30)
        print(f"{name}/{k}: z {np.round(z, 2)} (t crit {np.round(crit, 2)})", flush=True)
        tol = crit * np.sqrt(se2) + 1e-6 * (1 + np.abs(mr))
        assert np.all(np.abs(mg - mr) <= tol), f"{name}/{k}: gpu {mg} oracle {mr} tol {tol}"


def test_radiance_variance_map(vpt, oracle):
    """The Radiance Variance map (an extension) is E[x^2] - E[x]^2 of the accumulated samples: the same statistic
    computed from the running means of 1-spp frames rendered with the same random numbers."""
    print("case: radiance variance map delta 20^3", flush=True)
    torch = _torch()
    _setup(vpt, oracle.fbm_cloud(20, 20, 20, seed=33), 0)
    vpt.set_use_feature_maps(["Radiance Variance"])
    t = torch.zeros((4, RES, RES), dtype=torch.float32, device="cuda")
    frames = 6
    s1 = np.zeros((3, RES, RES), np.float64)
    s2 = np.zeros((3, RES, RES), np.float64)
    prev = np.zeros((3, RES, RES), np.float64)
    for k in range(frames):
        mean = vpt.render_frame(t, 1).cpu().numpy()[:3].astype(np.float64)
        x = (k + 1) * mean - k * prev
        prev = mean
        s1 += x
        s2 += x * x
    var = vpt.get_feature_map(t, 21).cpu().numpy()
    ref = np.maximum(s2 / frames - (s1 / frames) ** 2, 0.0)
    np.testing.assert_allclose(var[:3], ref, rtol=2e-4, atol=2e-5)
    np.testing.assert_array_equal(var[3], 1.0)
    vpt.set_seed_offset(0)
    one = vpt.render_frame(t, frames)  # the same samples in one call: the same map
    np.testing.assert_allclose(vpt.get_feature_map(t, 21).cpu().numpy(), var, rtol=1e-4, atol=1e-6)
    assert one is not None
