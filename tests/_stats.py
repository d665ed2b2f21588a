"""Statistics shared by the GPU parity tests.

Every statistical comparison is a two-sided z-test of a family of m comparisons at the family-wise error rate of a
single 3-sigma test (alpha = 0.0027), Bonferroni-corrected: |z| <= zcrit(m), so zcrit(1) = 3.0, zcrit(3) = 3.32 for
the three colour channels of an image mean. Every test uses fixed seeds, so its outcome is reproducible. Standard
errors come from per-pixel sample variances on both sides (the oracle's sum of squared path radiances; on the GPU
the Radiance Variance map of the same render), so they are estimated from thousands of pixels rather than from a
handful of repeated renders.
"""
import math

import numpy as np

FAMILY_ALPHA = 0.0027  # two-sided tail of one 3-sigma comparison


def zcrit(m: int = 1, alpha: float = FAMILY_ALPHA) -> float:
    """Critical |z| of each of m comparisons so that the family errs with probability alpha (Bonferroni)."""
    from scipy.stats import norm

    return float(norm.isf(alpha / (2.0 * m)))


def tcrit(m: int, dof: float, alpha: float = FAMILY_ALPHA) -> float:
    """zcrit for comparisons whose standard errors are estimated from few repeated batches: Student's t with dof
    degrees of freedom (Welch) at the same Bonferroni-corrected family-wise error."""
    from scipy.stats import t

    return float(t.isf(alpha / (2.0 * m), dof))


def gpu_image_stats(vpt, frames, res, device="cuda"):
    """Renders `frames` samples per pixel of the current scene at res x res (res: an int, or (width, height)); the
    accumulation starts from zero. Returns (per-channel image mean, its standard error, image [4, H, W]). The
    standard error is sqrt(sum over pixels of the unbiased per-pixel sample variance / frames) / pixels, the
    variance coming from the Radiance Variance map of the same render."""
    import torch

    w, h = (res, res) if np.ndim(res) == 0 else res
    vpt.set_use_feature_maps(["Radiance Variance"])
    try:
        t = torch.zeros((4, h, w), dtype=torch.float32, device=device)
        img = vpt.render_frame(t, frames).detach().cpu().numpy().astype(np.float64)
        var = vpt.get_feature_map(t, 21)[:3].detach().cpu().numpy().astype(np.float64)
    finally:
        vpt.set_use_feature_maps([])
    var *= frames / max(frames - 1, 1)
    npix = w * h
    mean = img[:3].reshape(3, -1).mean(1)
    se = np.sqrt(var.reshape(3, -1).sum(1) / frames) / npix
    return mean, se, img


def oracle_image_stats(scene, frames, frame_start=0, threads=0):
    """The oracle's image after `frames` samples (the exact FP64 sample mean of each pixel): (per-channel mean,
    standard error, image [H, W, 3])."""
    img64, sq = scene.render_mean(frames, frame_start=frame_start, threads=threads)
    npix = img64.shape[0] * img64.shape[1]
    var = np.maximum(sq / frames - img64 ** 2, 0.0) * (frames / max(frames - 1, 1))
    return img64.reshape(-1, 3).mean(0), np.sqrt(var.reshape(-1, 3).sum(0) / frames) / npix, img64


def assert_means_equal(m_a, s_a, m_b, s_b, what, m=None, rel_floor=1e-6):
    """|m_a - m_b| <= zcrit(m) * sqrt(s_a^2 + s_b^2) per component (m defaults to the number of components), plus a
    relative floor for float32 accumulation rounding. Prints the z-scores."""
    m_a, s_a, m_b, s_b = (np.atleast_1d(np.asarray(v, np.float64)) for v in (m_a, s_a, m_b, s_b))
    k = m if m is not None else m_a.size
    se = np.sqrt(s_a ** 2 + s_b ** 2)
    tol = zcrit(k) * se + rel_floor * np.maximum(np.abs(m_a), np.abs(m_b))
    z = (m_a - m_b) / np.maximum(se, 1e-30)
    print(f"{what}: a {m_a} b {m_b} z {np.round(z, 2)} (zcrit {zcrit(k):.2f})", flush=True)
    assert np.all(np.abs(m_a - m_b) <= tol), f"{what}: a {m_a} b {m_b} z {z} tol {tol}"


def z_scores_vs_exact(mean, m2, exact, n):
    """Per-ray z of a continuous transmittance estimate (mean and mean of squares over n samples) against the exact
    value; the variance is floored at 1/n^2 so rays without a single nonzero sample stay finite."""
    mean = np.asarray(mean, np.float64)
    var = np.maximum(np.asarray(m2, np.float64) - mean ** 2, 0.0) * (n / max(n - 1, 1))
    se = np.sqrt(np.maximum(var / n, 1.0 / (n * n)))
    return (mean - np.asarray(exact, np.float64)) / se


def assert_unbiased_rays(z, what):
    """A family of per-ray z-scores of continuous estimators: each within zcrit(rays), and their mean within
    zcrit(1) / sqrt(rays)."""
    z = np.asarray(z, np.float64)
    zc = zcrit(z.size)
    zm = float(z.mean()) * math.sqrt(z.size)
    print(f"{what}: max|z| {np.max(np.abs(z)):.2f} (zcrit {zc:.2f}), mean z * sqrt(n) {zm:.2f}", flush=True)
    assert np.max(np.abs(z)) <= zc, (what, int(np.argmax(np.abs(z))), float(np.max(np.abs(z))))
    assert abs(zm) <= zcrit(1), (what, zm)


def assert_unbiased_binary_rays(mean, m2, exact, n, what):
    """Per-ray estimates of a binary (0 / 1 per sample) transmittance estimator: the count of ones over n samples is
    Binomial(n, T) with T the exact transmittance, so each ray gets an exact two-sided binomial p-value (no normal
    approximation, which fails for the rays with n T of a few), Bonferroni-corrected over the rays; the mean of the
    exactly standardised errors (variance T (1 - T) / n) stays within zcrit(1) / sqrt(rays)."""
    from scipy.stats import binom

    mean = np.asarray(mean, np.float64)
    T = np.clip(np.asarray(exact, np.float64), 0.0, 1.0)
    np.testing.assert_allclose(np.asarray(m2, np.float64), mean, rtol=1e-6, atol=1e-7)  # binary samples
    k = np.rint(mean * n).astype(np.int64)
    p = 2.0 * np.minimum(binom.cdf(k, n, T), binom.sf(k - 1, n, T))
    p = np.minimum(p, 1.0)
    z = (mean - T) / np.sqrt(np.maximum(T * (1.0 - T), 1e-300) / n)
    zm = float(z.mean()) * math.sqrt(z.size)
    print(f"{what}: min p {p.min():.2e} (limit {FAMILY_ALPHA / p.size:.2e}), mean z * sqrt(n) {zm:.2f}", flush=True)
    assert p.min() >= FAMILY_ALPHA / p.size, (what, int(np.argmin(p)), float(p.min()), int(k[np.argmin(p)]),
                                               float(T[np.argmin(p)]))
    assert abs(zm) <= zcrit(1), (what, zm)


def assert_converged(gpu, var_gpu, n_gpu, ref, var_ref, n_ref, what, rel_max=None):
    """Two converged renders [H, W, 3] with their per-pixel sample variances and sample counts. Values with sampling
    noise (standard error above 1e-4 of the value) get z = (gpu - ref) / sqrt(var_gpu / n_gpu + var_ref / n_ref): a
    bias shows as a shifted mean of z, so per channel mean z * sqrt(values) stays within zcrit(3). mean(z^2) is 1 for
    exact variances and somewhat above when they are estimated from heavy-tailed samples (next-event estimation:
    1.5 measured), so it is only required to stay below 2; a misregistered view or a pixel-dependent error lifts it
    far above. Values without sampling noise (pixels that see only the sky) must agree to 1e-3 relative (float
    evaluation of the sky on each side). Returns (relMSE, z of the noisy values)."""
    gpu, ref = np.asarray(gpu, np.float64), np.asarray(ref, np.float64)
    se = np.sqrt(np.asarray(var_gpu, np.float64) / n_gpu + np.asarray(var_ref, np.float64) / n_ref)
    noisy = se > 1e-4 * np.abs(ref)
    assert np.all(np.abs(gpu - ref)[~noisy] <= 1e-3 * np.abs(ref[~noisy]) + 1e-6), (what, "noise-free values")
    z = np.where(noisy, (gpu - ref) / np.maximum(se, 1e-30), 0.0)
    cnt = noisy.reshape(-1, 3).sum(0)
    zbar = z.reshape(-1, 3).sum(0) / np.sqrt(np.maximum(cnt, 1))
    z2 = float((z * z).sum() / max(int(noisy.sum()), 1))
    rel = float(np.mean((gpu - ref) ** 2 / (ref * ref + 0.01)))
    print(f"{what}: relMSE {rel:.3e}, {int(noisy.sum())} noisy values, mean z * sqrt(values) {np.round(zbar, 2)} "
          f"(zcrit {zcrit(3):.2f}), mean z^2 {z2:.3f}", flush=True)
    assert np.all(np.abs(zbar) <= zcrit(3)), (what, zbar)
    assert z2 <= 2.0, (what, z2)
    if rel_max is not None:
        assert rel <= rel_max, (what, rel)
    return rel, z
