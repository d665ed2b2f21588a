"""CPU pins of the KPN apply restatement: one-hot kernels shift the image (zero where the tap leaves the output
extent), the centre tap is the identity, the operator is linear in the weights, and a zero-logit gate blends the
inputs half and half."""
import numpy as np

from oracle import kpn_ref as K


def test_one_hot_kernels_shift_the_image():
    rng = np.random.default_rng(0)
    img = rng.random((2, 3, 9, 11)).astype(np.float32)
    k = 5
    for idx in range(k * k):
        dy, dx = idx // k - 2, idx % k - 2
        w = np.zeros((2, k * k, 9, 11), np.float32)
        w[:, idx] = 1.0
        out = K.per_pixel_kernel(img, w, k)
        ref = np.zeros_like(img)
        ys, xs = np.mgrid[0:9, 0:11]
        ok = (ys + dy >= 0) & (ys + dy < 9) & (xs + dx >= 0) & (xs + dx < 11)
        ref[:, :, ok] = img[:, :, (ys + dy)[ok], (xs + dx)[ok]]
        np.testing.assert_array_equal(out, ref)


def test_linearity_and_smaller_weight_extent():
    rng = np.random.default_rng(1)
    img = rng.random((1, 4, 10, 12)).astype(np.float32)
    w1 = rng.random((1, 9, 8, 7)).astype(np.float32)
    w2 = rng.random((1, 9, 8, 7)).astype(np.float32)
    a = K.per_pixel_kernel(img, w1, 3) + K.per_pixel_kernel(img, w2, 3)
    b = K.per_pixel_kernel(img, w1 + w2, 3)
    assert a.shape == (1, 4, 8, 7)
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-6)


def test_zero_gates_blend_evenly():
    rng = np.random.default_rng(2)
    x = rng.random((1, 3, 8, 8)).astype(np.float32)
    lo_x = rng.random((1, 3, 4, 4)).astype(np.float32)
    lo_p = rng.random((1, 3, 4, 4)).astype(np.float32)
    prev = rng.random((1, 3, 8, 8)).astype(np.float32)
    fused = np.zeros((1, 9 + 3, 8, 8), np.float32)
    base = K.inv_iter_base(x, lo_p, lo_x, prev, fused, 3, True)
    up = lambda a: np.repeat(np.repeat(a, 2, 2), 2, 3)
    lf = 0.5 * (0.5 * up(lo_p) + 0.5 * up(lo_x)) + 0.5 * prev
    ref = 0.5 * (x - up(lo_x) + lf) + 0.5 * lf
    np.testing.assert_allclose(base, ref, rtol=1e-6, atol=1e-6)
