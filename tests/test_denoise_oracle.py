"""CPU pins of the denoiser restatement (oracle/denoise_ref.py) before it checks the gfx950 kernels.

EAW: with only the colour weight (the reference's constant exp(-1/phiColor)) each iteration is a plain dilated 5x5
Gaussian (weights exp(-(x^2+y^2)/2)) with clamped edges, so it must equal scipy's correlation; zero iterations are
the identity. SVGF: every filter is normalised, so a constant image stays constant; depth and normal edges stop the
filters (a two-region image keeps each region's value away from the edge); a static camera accumulates history
(the second frame's temporal weight is 1/2).
"""
import numpy as np
import pytest
from scipy import ndimage

from oracle import denoise_ref as D


def _rng_img(h, w, seed):
    return np.random.default_rng(seed).random((h, w, 4)).astype(np.float32)


@pytest.mark.parametrize("iterations", [1, 2, 3])
def test_eaw_colour_only_is_dilated_gaussian(iterations):
    img = _rng_img(23, 31, 1)
    got = D.eaw(img, iterations=iterations, phi_color=5.0)
    ref = img.astype(np.float64)
    for i in range(iterations):
        step = 1 << i
        k = np.zeros((4 * step + 1, 4 * step + 1))
        for y in range(-2, 3):
            for x in range(-2, 3):
                k[2 * step + y * step, 2 * step + x * step] = np.exp(-(x * x + y * y) / 2.0)
        k /= k.sum()
        ref = np.stack([ndimage.correlate(ref[..., c], k, mode="nearest") for c in range(4)], -1)
    np.testing.assert_allclose(got, ref, rtol=2e-5, atol=2e-6)


def test_eaw_zero_iterations_identity_and_weights_cut_edges():
    img = _rng_img(16, 16, 2)
    np.testing.assert_array_equal(D.eaw(img, iterations=0), img)
    # two position regions far apart: no pixel of the left half mixes with the right half
    pos = np.zeros((16, 16, 4), np.float32)
    pos[:, 8:, 0] = 100.0
    col = np.zeros((16, 16, 4), np.float32)
    col[:, 8:] = 1.0
    out = D.eaw(col, position=pos, iterations=2)
    np.testing.assert_allclose(out[:, :8], 0.0, atol=1e-6)
    np.testing.assert_allclose(out[:, 8:], 1.0, atol=1e-6)


def _inputs(h, w, seed, depth=1.0):
    rng = np.random.default_rng(seed)
    noisy = rng.random((h, w, 4)).astype(np.float32)
    motion = np.zeros((h, w, 2), np.float32)
    normal = np.zeros((h, w, 4), np.float32)
    normal[..., 2] = 1.0
    normal[..., 3] = 1.0
    dep = np.full((h, w), depth, np.float32)
    fw = np.full((h, w), 0.01, np.float32)
    return noisy, motion, normal, dep, fw


def test_svgf_constant_image_stays_constant():
    h, w = 20, 24
    noisy, motion, normal, dep, fw = _inputs(h, w, 3)
    noisy[..., :3] = 0.37
    s = D.Svgf(h, w)
    for _ in range(3):
        out = s.denoise(noisy, motion, normal, dep, fw)
        np.testing.assert_allclose(out[..., :3], 0.37, rtol=1e-5)


def test_svgf_depth_and_normal_edges_stop_the_filter():
    h, w = 24, 24
    noisy, motion, normal, dep, fw = _inputs(h, w, 4)
    noisy[..., :3] = 0.0
    noisy[:, 12:, :3] = 1.0
    dep[:, 12:] = 5.0   # a depth edge between the halves
    normal[12:, :, :3] = [1.0, 0.0, 0.0]  # a normal edge between the top and bottom halves
    s = D.Svgf(h, w)
    out = s.denoise(noisy, motion, normal, dep, fw)
    np.testing.assert_allclose(out[:, :12, :3], 0.0, atol=1e-6)
    np.testing.assert_allclose(out[:, 12:, :3], 1.0, atol=1e-6)


def test_svgf_static_history_accumulates():
    """Frame 1 has no valid history (depths differ from the zero history); in frame 2 every interior pixel reprojects
    onto itself with history length 2, so its temporal colour is the mean of this frame's colour and the history
    fetched by the 2x2 bilinear tap at the reference's +0.01 offset (weights 0.99^2, 0.99*0.01, 0.01*0.99, 0.01^2)."""
    h, w = 12, 12
    a, motion, normal, dep, fw = _inputs(h, w, 5)
    b = _inputs(h, w, 6)[0]
    s = D.Svgf(h, w, iterations=0)
    s.denoise(a, motion, normal, dep, fw)
    hist = s.color_history[..., :3].astype(np.float64)
    acc, temp = s.reproject(b, motion, normal, dep)
    inner = (slice(1, h - 1), slice(1, w - 1))
    np.testing.assert_array_equal(acc[inner][..., 2], 2.0)
    f = 0.01
    lerp = ((1 - f) ** 2 * hist[1:h - 1, 1:w - 1] + (1 - f) * f * hist[2:h, 1:w - 1] +
            f * (1 - f) * hist[1:h - 1, 2:w] + f * f * hist[2:h, 2:w])
    np.testing.assert_allclose(temp[inner][..., :3], 0.5 * (lerp + b[inner][..., :3]), rtol=2e-5)
