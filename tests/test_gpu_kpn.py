"""GPU KPN apply kernels against the NumPy restatement (oracle/kpn_ref.py) on random tensors."""
import numpy as np
import pytest

from oracle import kpn_ref as K

pytestmark = pytest.mark.gpu


def _t(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.mark.parametrize("k,n,c", [(5, 2, 3), (3, 1, 4), (5, 1, 1)])
def test_per_pixel_kernel(vpt, k, n, c):
    print(f"case: per-pixel kernel k={k} n={n} c={c}", flush=True)
    rng = np.random.default_rng(k * 10 + n + c)
    img = rng.random((n, c, 33, 47)).astype(np.float32)
    w = rng.random((n, k * k, 31, 45)).astype(np.float32)
    got = vpt.per_pixel_kernel_forward(_t(img), _t(w), k).cpu().numpy()
    np.testing.assert_allclose(got, K.per_pixel_kernel(img, w, k), rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("use_previous,lo", [(False, (8, 10)), (True, (8, 10)), (True, (5, 7))])
def test_inv_iter_kernel(vpt, use_previous, lo):
    """lo (5, 7) leaves pixels with x // 2 >= 7 or y // 2 >= 5 outside the low-resolution image: their base is the
    reference's zero-initialised result (Kernel.cu:240), whatever the scratch held before."""
    print(f"case: inv iter kernel previous={use_previous} lo={lo}", flush=True)
    rng = np.random.default_rng(7)
    n, c, k = 2, 3, 5
    x = rng.random((n, c, 16, 20)).astype(np.float32)
    lo_x = rng.random((n, c) + lo).astype(np.float32)
    lo_p = rng.random((n, c) + lo).astype(np.float32)
    prev = rng.random((n, c, 16, 20)).astype(np.float32)
    fused = rng.normal(size=(n, k * k + 3, 16, 20)).astype(np.float32)
    got = vpt.inv_iter_kernel(_t(x), _t(lo_p), _t(lo_x), _t(prev), _t(fused), k, use_previous).cpu().numpy()
    ref = K.inv_iter_kernel(x, lo_p, lo_x, prev, fused, k, use_previous)
    np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-5)
