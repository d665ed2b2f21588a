"""GPU EAW and SVGF denoisers against the NumPy restatement (oracle/denoise_ref.py), fed with the feature maps the
renderer produced for the same frame."""
import numpy as np
import pytest

from oracle import denoise_ref as D

pytestmark = pytest.mark.gpu

RES = 32


def _torch():
    import torch

    return torch


def _hwc(t):
    return np.ascontiguousarray(np.moveaxis(t.cpu().numpy(), 0, -1))


def _scene(vpt, oracle):
    vpt.set_density_field(oracle.fbm_cloud(24, 20, 16, seed=41))
    vpt.set_vpt_mode(0)
    vpt.set_grid_interpolation_type("stochastic")
    vpt.set_extinction_scale(80.0)
    vpt.set_extinction_base([1.0, 1.0, 1.0])
    vpt.set_scattering_albedo([0.9] * 3)
    vpt.set_phase_g(0.5)
    vpt.set_use_builtin_environment_map("Default")
    vpt.set_use_lights(False)
    vpt.set_seed_offset(0)
    vpt.set_use_feature_maps([])


@pytest.mark.parametrize("iterations", [0, 1, 3])
def test_eaw_matches_restatement(vpt, oracle, iterations):
    print(f"case: EAW {iterations} iterations", flush=True)
    torch = _torch()
    _scene(vpt, oracle)
    vpt.set_denoiser("EAW")
    vpt.set_denoiser_property("maxNumIterations", str(iterations))
    vpt.set_use_feature_maps(["First X", "Normal"])
    t = torch.zeros((4, RES, RES), dtype=torch.float32, device="cuda")
    out = _hwc(vpt.render_frame(t, 8))
    res = _hwc(vpt.get_feature_map_from_string(t, "Result"))
    pos = _hwc(vpt.get_feature_map_from_string(t, "First X"))
    nrm = _hwc(vpt.get_feature_map_from_string(t, "Normal"))
    vpt.set_denoiser("None")
    ref = D.eaw(res, pos, nrm, iterations=iterations)
    np.testing.assert_allclose(out, ref, rtol=2e-4, atol=2e-6)


def test_eaw_odd_size_all_step_widths(vpt, oracle):
    """A 44 x 37 frame through five passes (step widths 1 to 16): the passes stage each phase of the dilated lattice
    in LDS tiles, so widths that do not divide the image and taps clamped at every edge are covered."""
    print("case: EAW 44x37, 5 iterations", flush=True)
    torch = _torch()
    _scene(vpt, oracle)
    vpt.set_denoiser("EAW")
    vpt.set_denoiser_property("maxNumIterations", "5")
    vpt.set_use_feature_maps(["First X", "Normal"])
    t = torch.zeros((4, 37, 44), dtype=torch.float32, device="cuda")
    out = _hwc(vpt.render_frame(t, 8))
    res = _hwc(vpt.get_feature_map_from_string(t, "Result"))
    pos = _hwc(vpt.get_feature_map_from_string(t, "First X"))
    nrm = _hwc(vpt.get_feature_map_from_string(t, "Normal"))
    vpt.set_denoiser("None")
    vpt.set_use_feature_maps([])
    ref = D.eaw(res, pos, nrm, iterations=5)
    np.testing.assert_allclose(out, ref, rtol=2e-4, atol=2e-6)


def test_svgf_matches_restatement_over_frames(vpt, oracle):
    print("case: SVGF 4 frames", flush=True)
    torch = _torch()
    _scene(vpt, oracle)
    vpt.set_denoiser("SVGF")
    s = D.Svgf(RES, RES)
    t = torch.zeros((4, RES, RES), dtype=torch.float32, device="cuda")
    try:
        for frame in range(4):
            out = _hwc(vpt.render_frame(t, 1))
            m = {n: _hwc(vpt.get_feature_map_from_string(t, n))
                 for n in ["Result", "Flow", "Normal", "Depth", "Depth (fwidth)"]}
            ref = s.denoise(m["Result"], m["Flow"][..., :2], m["Normal"], m["Depth"][..., 0],
                            m["Depth (fwidth)"][..., 0])
            assert np.all(np.isfinite(out))
            np.testing.assert_allclose(out[..., :3], ref[..., :3], rtol=2e-3, atol=2e-5,
                                       err_msg=f"frame {frame}")
    finally:
        vpt.set_denoiser("None")
