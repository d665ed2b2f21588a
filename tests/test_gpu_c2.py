"""Transmittance parity on the headline sparse volume: 10^6 random rays per estimator against the FP64 oracle.

Volume: NanoVDB FP32 grid of a 512^3 seeded fBm (seed 0, smoothstep(0.645, 0.795), about 12 % active voxels),
extinction 1024, stochastic filtering (its expectation is the trilinear field, which the FP64 integrator evaluates
exactly per voxel with 2-point Gauss-Legendre). Rays: origins uniform on a sphere of radius 1.5 x the box's
half-diagonal around the box centre, directions towards uniform points in the box, seed 1. One estimate per ray:
the mean error over all rays, and within each of the six bins of the exact transmittance, must be within zcrit(7) =
3.78 standard errors (the family-wise 3-sigma level of the seven comparisons).
"""
import math

import numpy as np
import pytest

from _stats import zcrit

pytestmark = pytest.mark.gpu

N_RAYS = 1_000_000
GRID = 512


@pytest.fixture(scope="module")
def c2(vpt, oracle):
    print("case: c2 volume 512^3 and 10^6 FP64 optical depths", flush=True)
    vpt.generate_fbm_cloud(GRID, GRID, GRID, seed=0, edge0=0.645, edge1=0.795)
    vpt.set_grid_interpolation_type("stochastic")
    vpt.set_extinction_base([1.0, 1.0, 1.0])
    vpt.set_extinction_scale(1024.0)
    vpt.set_use_sparse_grid(True)
    box = vpt.get_render_bounding_box()
    bmin = np.array(box[0::2], np.float64)
    bmax = np.array(box[1::2], np.float64)
    d = vpt.get_density_field((GRID, GRID, GRID), normalized=True)
    active = float(np.count_nonzero(d)) / d.size
    nzm = d != 0
    zs, ys, xs = (np.nonzero(nzm.any(axis=ax))[0] for ax in ((1, 2), (0, 2), (0, 1)))
    del nzm
    crop = np.ascontiguousarray(d[zs[0]:zs[-1] + 1, ys[0]:ys[-1] + 1, xs[0]:xs[-1] + 1])
    del d
    sc = oracle.Scene(crop, interpolation="stochastic", extinction=1024.0, voxel_range=(0.0, 1.0),
                      box=(bmin.astype(np.float32), bmax.astype(np.float32)))
    rng = np.random.default_rng(1)
    c = 0.5 * (bmin + bmax)
    radius = 1.5 * np.linalg.norm(0.5 * (bmax - bmin))
    o = rng.normal(size=(N_RAYS, 3))
    o = c + o / np.linalg.norm(o, axis=1, keepdims=True) * radius
    tgt = rng.uniform(bmin, bmax, (N_RAYS, 3))
    w = tgt - o
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    o = o.astype(np.float32)
    w = w.astype(np.float32)
    exact = np.exp(-sc.optical_depth(o, w, threads=16))
    print(f"c2 volume: {active:.4f} active, exact T mean {exact.mean():.4f}", flush=True)
    return {"o": o, "w": w, "exact": exact, "active": active}


def test_c2_active_fraction(c2):
    assert 0.09 <= c2["active"] <= 0.15, c2["active"]


@pytest.mark.parametrize("estimator", ["delta", "ratio", "residual_ratio", "decomposition"])
def test_c2_transmittance_1e6_rays_vs_fp64(vpt, c2, estimator):
    print(f"case: c2 transmittance {estimator} 10^6 rays", flush=True)
    vpt.set_use_sparse_grid(True)
    mean, _ = vpt.transmittance_batch(estimator, c2["o"], c2["w"], n_samples=1, seed=1)
    exact = c2["exact"]
    diff = mean.astype(np.float64) - exact
    se = diff.std() / math.sqrt(len(diff))
    z = diff.mean() / max(se, 1e-12)
    zc = zcrit(7)
    print(f"c2 {estimator}: mean error {diff.mean():.3e} z {z:.2f} (zcrit {zc:.2f})", flush=True)
    assert abs(z) <= zc, f"{estimator}: mean error {diff.mean():.3e}, z {z:.2f}"
    edges = [0.0, 0.05, 0.25, 0.5, 0.75, 0.95, 1.0 + 1e-9]
    for lo, hi in zip(edges[:-1], edges[1:]):
        sel = (exact >= lo) & (exact < hi)
        if sel.sum() < 1000:
            continue
        db = diff[sel]
        zb = db.mean() / max(db.std() / math.sqrt(len(db)), 1e-12)
        print(f"  bin [{lo}, {hi}) n={sel.sum()}: z {zb:.2f}", flush=True)
        assert abs(zb) <= zc, f"{estimator}: bin [{lo}, {hi}) n={sel.sum()} mean error {db.mean():.3e} z {zb:.2f}"
