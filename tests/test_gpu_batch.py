"""GPU tests of the batched render queue (an extension without a reference counterpart).

A request draws exactly the random numbers of a single-frame render of the same scene (pixel, frame number, seed
offset), and the single-frame path kernel and the batch kernel run the same device code, so every path takes the
same decisions and every request's image equals render_frame's up to the last bits: the compiler forms fused
multiply-adds differently in the two kernels (measured: at most a few ulp on 4-15 % of the pixels). A decision
whose random number lies within an ulp of its threshold could still flip in rare cases, so up to 0.2 % of the
pixels may differ by more. With several samples per unit the float sum over samples is regrouped as well. The
single-frame renderer's parity against the oracle is covered by the other GPU tests.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


# (volume, (nx, ny, nz), seed, sparse, camera position, g, seed offset, (channels, h, w), spp)
SCENES = [
    ("dense", (32, 28, 24), 3, False, (0.0, 0.0, 0.8), 0.5, 0, (3, 20, 24), 5),
    ("dense", (32, 28, 24), 3, False, (0.3, 0.2, 0.7), 0.85, 11, (4, 16, 16), 3),
    ("sparse", (40, 40, 40), 4, True, (0.0, 0.1, 0.8), 0.0, 5, (3, 24, 32), 7),
    ("dense", (20, 20, 20), 9, False, (-0.2, 0.0, 0.75), 0.3, 2, (3, 12, 12), 2),
]


def _set_scene(vpt, sc, mode, last_volume):
    name, dims, seed, sparse, cam, g, offset, _, _ = sc
    key = (dims, seed, sparse)
    if key != last_volume:
        vpt.generate_fbm_cloud(*dims, seed=seed, edge0=0.45, edge1=0.6)
        vpt.set_use_sparse_grid(sparse)
    vpt.set_vpt_mode(mode)
    vpt.set_grid_interpolation_type("stochastic")
    vpt.set_extinction_scale(96.0)
    vpt.set_extinction_base([1.0, 1.0, 1.0])
    vpt.set_scattering_albedo([0.9, 0.9, 0.9])
    vpt.set_use_builtin_environment_map("Default")
    vpt.set_use_lights(mode == 5)
    vpt.set_camera_position(list(cam))
    vpt.set_camera_target([0.0, 0.0, 0.0])
    vpt.set_phase_g(g)
    vpt.set_seed_offset(offset)
    vpt.set_max_bounces(32)
    return key


def _queue(vpt, mode, scenes=SCENES):
    torch = _torch()
    vpt.batch_clear()
    last = None
    for sc in scenes:
        last = _set_scene(vpt, sc, mode, last)
        c, h, w = sc[7]
        vpt.batch_add_request(torch.zeros((c, h, w), dtype=torch.float32, device="cuda"), sc[8])


def _assert_same_paths(a, b, what):
    """Equal up to the rounding of the two kernels' arithmetic: 2e-6 relative (a few ulp) on at least 99.8 % of the
    values, image means within 1e-6 relative."""
    close = np.isclose(a, b, rtol=2e-6, atol=2e-7)
    assert close.mean() >= 0.998, f"{what}: {int((~close).sum())} of {close.size} values differ"
    assert abs(float(a.mean()) - float(b.mean())) <= 1e-6 * max(abs(float(b.mean())), 1e-3), what


def _references(vpt, mode, scenes=SCENES):
    torch = _torch()
    refs = []
    last = None
    for sc in scenes:
        last = _set_scene(vpt, sc, mode, last)
        c, h, w = sc[7]
        refs.append(vpt.render_frame(torch.zeros((c, h, w), dtype=torch.float32, device="cuda"), sc[8]).cpu().numpy())
    return refs


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4, 5, 6])
def test_batch_matches_single_frame_renders(vpt, mode):
    print(f"case: batch of {len(SCENES)} requests, mode {mode}, one sample per unit", flush=True)
    torch = _torch()
    _queue(vpt, mode)
    vpt.batch_set_launch_config(0, 256, 8, 1)
    outs = vpt.batch_render()
    torch.cuda.synchronize()
    got = [o.cpu().numpy() for o in outs]
    refs = _references(vpt, mode)
    for i, (a, b) in enumerate(zip(got, refs)):
        assert np.isfinite(a).all(), f"request {i}"
        _assert_same_paths(a, b, f"request {i}")
    st = vpt.batch_get_stats()
    assert st["num_requests"] == len(SCENES)
    assert st["num_paths"] == sum(s[7][1] * s[7][2] * s[8] for s in SCENES)
    vpt.batch_clear()


@pytest.mark.parametrize("spu,block,steps", [(4, 64, 2), (16, 128, 16), (3, 192, 5)])
def test_batch_samples_per_unit_regroups_only(vpt, spu, block, steps):
    print(f"case: batch samples per unit {spu}, block {block}, steps {steps}", flush=True)
    torch = _torch()
    _queue(vpt, 3)
    vpt.batch_set_launch_config(0, block, steps, spu)
    outs = vpt.batch_render()
    torch.cuda.synchronize()
    refs = _references(vpt, 3)
    for i, (a, b) in enumerate(zip(outs, refs)):
        _assert_same_paths(a.cpu().numpy(), b, f"request {i}")
    vpt.batch_clear()


def test_batch_graph_replay_and_frame_offset(vpt):
    print("case: batch graph replays and frame offsets", flush=True)
    torch = _torch()
    scenes = SCENES[:3]
    _queue(vpt, 0, scenes)
    vpt.batch_set_launch_config(0, 256, 8, 1)
    first = [o.clone() for o in vpt.batch_render(0)]
    for _ in range(3):
        torch.cuda.synchronize()
        again = vpt.batch_render(0)
        torch.cuda.synchronize()
        for a, b in zip(first, again):
            assert torch.equal(a, b)
    # frames spp .. 2 spp - 1 are the second render_frame call of an accumulating sequence
    shifted = [o.clone() for o in vpt.batch_render(scenes[0][8])]
    torch.cuda.synchronize()
    assert not torch.equal(first[0], shifted[0])
    sc = scenes[0]
    _set_scene(vpt, sc, 0, None)
    c, h, w = sc[7]
    vpt.render_frame(torch.zeros((c, h, w), dtype=torch.float32, device="cuda"), sc[8])
    two = vpt.render_frame(torch.zeros((c, h, w), dtype=torch.float32, device="cuda"), sc[8])
    mean = 0.5 * (first[0] + shifted[0])
    np.testing.assert_allclose(mean.cpu().numpy(), two.cpu().numpy(), rtol=1e-5, atol=1e-6)
    vpt.batch_clear()


def test_batch_autotune(vpt):
    print("case: batch autotune", flush=True)
    torch = _torch()
    _queue(vpt, 0)
    cfg = vpt.batch_autotune(2)
    assert cfg["block_size"] in (64, 128, 256)
    assert cfg["steps_per_refill"] > 0 and cfg["samples_per_unit"] > 0 and cfg["blocks"] > 0
    assert vpt.batch_get_launch_config() == cfg
    outs = vpt.batch_render()
    torch.cuda.synchronize()
    refs = _references(vpt, 0)
    for i, (a, b) in enumerate(zip(outs, refs)):
        _assert_same_paths(a.cpu().numpy(), b, f"request {i}")
    vpt.batch_clear()


def test_batch_rejects_mixed_estimators_and_bad_configs(vpt):
    print("case: batch argument errors", flush=True)
    torch = _torch()
    _queue(vpt, 0, SCENES[:1])
    vpt.set_vpt_mode(2)
    with pytest.raises(vpt.CrvptError):
        vpt.batch_add_request(torch.zeros((3, 8, 8), dtype=torch.float32, device="cuda"), 4)
    with pytest.raises(vpt.CrvptError):
        vpt.batch_set_launch_config(0, 96, 8, 1)
    with pytest.raises(vpt.CrvptError):
        vpt.batch_set_launch_config(0, 256, 0, 1)
    vpt.batch_clear()
    with pytest.raises(vpt.CrvptError):
        vpt.batch_render()
