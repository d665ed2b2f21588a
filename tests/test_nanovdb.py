"""CPU checks of the build's NanoVDB writer against the reference's own NanoVDB 32.3.3.

Golden fixtures (tests/golden/nanovdb_*.npz) hold grids written by the reference GridBuilder and lookups made by the
reference PNanoVDB accessor (tests/golden/make_nanovdb_golden.py, oracle/_ref/nvdb_ref). The build's writer must
reproduce the reference buffer byte for byte; when the reference tool is present (this container), random grids are
compared too.
"""
import glob
import os
import subprocess
import tempfile

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = sorted(p for p in glob.glob(os.path.join(REPO, "tests", "golden", "nanovdb_*.npz")) if "fp16" not in p)
GOLDEN16 = sorted(glob.glob(os.path.join(REPO, "tests", "golden", "nanovdb_fp16_*.npz")))
TOOL = os.path.join(REPO, "oracle", "_ref", "nvdb_ref")


def _build(dense, dx, origin, value_type=0):
    from cloudrendering_amd import vpt

    return np.frombuffer(vpt.build_nanovdb_from_dense(dense, dx, tuple(origin), value_type), np.uint8)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_writer_bit_exact_vs_reference_golden(path):
    g = np.load(path)
    mine = _build(g["dense"], float(g["dx"]), g["origin"])
    ref = g["grid"]
    assert mine.size == ref.size
    diff = np.nonzero(mine != ref)[0]
    assert diff.size == 0, f"first differing byte offsets {diff[:16]}"


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_reference_lookups_match_dense_input(path):
    """The reference accessor returns the dense value inside the grid and the background 0 outside: pins that the
    golden lookups and the dense data describe the same grid (x, y, z index order)."""
    g = np.load(path)
    dense, ijk, vals = g["dense"], g["ijk"], g["values"]
    nz, ny, nx = dense.shape
    inside = (ijk[:, 0] >= 0) & (ijk[:, 1] >= 0) & (ijk[:, 2] >= 0) & (ijk[:, 0] < nx) & (ijk[:, 1] < ny) & \
        (ijk[:, 2] < nz)
    expect = np.zeros(len(ijk), np.float32)
    expect[inside] = dense[ijk[inside, 2], ijk[inside, 1], ijk[inside, 0]]
    np.testing.assert_array_equal(vals, expect)


@pytest.mark.skipif(not os.path.exists(TOOL), reason="reference NanoVDB tool not built (needs the reference tree)")
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_writer_bit_exact_vs_reference_tool_random(seed):
    rng = np.random.default_rng(seed)
    shape = tuple(int(v) for v in rng.integers(5, 70, 3))
    dense = rng.random(shape).astype(np.float32)
    dense[dense < 0.6] = 0.0
    if seed == 3:
        dense[:, :, :] = 0.0
        dense[2:, :, :] = 0.75  # large constant regions: leaf tiles
    dx = float(rng.uniform(0.001, 0.05))
    origin = tuple(float(v) for v in rng.uniform(-1, 1, 3))
    nz, ny, nx = shape
    with tempfile.TemporaryDirectory() as tmp:
        dense.tofile(os.path.join(tmp, "d.f32"))
        subprocess.run([TOOL, "build", os.path.join(tmp, "d.f32"), str(nx), str(ny), str(nz), repr(dx),
                        *[repr(v) for v in origin], os.path.join(tmp, "g.bin")], check=True)
        ref = np.fromfile(os.path.join(tmp, "g.bin"), np.uint8)
    mine = _build(dense, dx, origin)
    assert mine.size == ref.size
    assert np.array_equal(mine, ref), np.nonzero(mine != ref)[0][:16]


@pytest.mark.parametrize("path", GOLDEN16, ids=[os.path.basename(p) for p in GOLDEN16])
def test_fp16_writer_bit_exact_vs_reference_golden(path):
    """Fp16 leaves (GridBuilder<float, Fp16>): codes, per-leaf minimum and quantum, quantised statistics, the
    internal nodes' statistics over the decoded values, grid type and checksum, byte for byte."""
    g = np.load(path)
    mine = _build(g["dense"], float(g["dx"]), g["origin"], value_type=1)
    ref = g["grid"]
    assert mine.size == ref.size
    diff = np.nonzero(mine != ref)[0]
    assert diff.size == 0, f"first differing byte offsets {diff[:16]}"


@pytest.mark.parametrize("path", GOLDEN16, ids=[os.path.basename(p) for p in GOLDEN16])
def test_fp16_reference_lookups_quantise_the_dense_input(path):
    """The reference Fp16 accessor returns each dense value to within half a quantum of its leaf (quantum = leaf
    range / 65535, leaf range <= 1 for normalised inputs), exactly 0 outside the grid, and exactly the tile value
    where a constant block was pruned: pins the golden lookups against the mathematics of the format."""
    g = np.load(path)
    dense, ijk, vals = g["dense"], g["ijk"], g["values"]
    nz, ny, nx = dense.shape
    inside = (ijk[:, 0] >= 0) & (ijk[:, 1] >= 0) & (ijk[:, 2] >= 0) & (ijk[:, 0] < nx) & (ijk[:, 1] < ny) & \
        (ijk[:, 2] < nz)
    np.testing.assert_array_equal(vals[~inside], 0.0)
    ref = dense[ijk[inside, 2], ijk[inside, 1], ijk[inside, 0]]
    np.testing.assert_allclose(vals[inside], ref, rtol=0, atol=0.5 / 65535 + 1e-7)


@pytest.mark.skipif(not os.path.exists(TOOL), reason="reference NanoVDB tool not built (needs the reference tree)")
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_fp16_writer_bit_exact_vs_reference_tool_random(seed):
    rng = np.random.default_rng(100 + seed)
    shape = tuple(int(v) for v in rng.integers(5, 70, 3))
    dense = (rng.random(shape) * rng.uniform(0.1, 3.0)).astype(np.float32)
    dense[dense < 0.5 * dense.max()] = 0.0
    dx = float(rng.uniform(0.001, 0.05))
    origin = tuple(float(v) for v in rng.uniform(-1, 1, 3))
    nz, ny, nx = shape
    with tempfile.TemporaryDirectory() as tmp:
        dense.tofile(os.path.join(tmp, "d.f32"))
        subprocess.run([TOOL, "build16", os.path.join(tmp, "d.f32"), str(nx), str(ny), str(nz), repr(dx),
                        *[repr(v) for v in origin], os.path.join(tmp, "g.bin")], check=True)
        ref = np.fromfile(os.path.join(tmp, "g.bin"), np.uint8)
    mine = _build(dense, dx, origin, value_type=1)
    assert mine.size == ref.size
    assert np.array_equal(mine, ref), np.nonzero(mine != ref)[0][:16]


def test_fog_sphere_bit_exact_vs_reference_default_scene():
    """The default scene's grid (Module.cpp:793-795, nanovdb::createFogVolumeSphere<float>(0.25, 0, 0.01)) equals the
    reference's own createFogVolumeSphere output byte for byte (tests/golden/fog_sphere_default.npz)."""
    from cloudrendering_amd import vpt

    g = np.load(os.path.join(REPO, "tests", "golden", "fog_sphere_default.npz"))
    mine = np.frombuffer(vpt.build_fog_volume_sphere(0.25, (0.0, 0.0, 0.0), 0.01), np.uint8)
    ref = g["grid"]
    assert mine.size == ref.size
    diff = np.nonzero(mine != ref)[0]
    assert diff.size == 0, f"first differing byte offsets {diff[:16]}"


def test_fog_sphere_golden_lookups_are_the_fog_profile():
    """The golden lookups describe the fog sphere: 1 deeper than three voxels inside radius 25 voxels, the linear
    ramp -d / 3 across the band, 0 outside (pins that the fixture's lookups and the geometry agree)."""
    g = np.load(os.path.join(REPO, "tests", "golden", "fog_sphere_default.npz"))
    ijk, vals = g["ijk"].astype(np.float64), g["values"].astype(np.float64)
    d = np.sqrt((ijk ** 2).sum(1)) - 25.0
    expect = np.clip(-d / 3.0, 0.0, 1.0)
    np.testing.assert_allclose(vals, expect, atol=2e-6)
    assert np.any(vals == 1.0) and np.any((vals > 0) & (vals < 1))


@pytest.mark.skipif(not os.path.exists(TOOL), reason="reference NanoVDB tool not built (needs the reference tree)")
@pytest.mark.parametrize("radius,voxel", [(0.25, 0.02), (0.1, 0.005), (0.03, 0.01)])
def test_fog_sphere_bit_exact_vs_reference_tool(radius, voxel):
    from cloudrendering_amd import vpt

    mine = np.frombuffer(vpt.build_fog_volume_sphere(radius, (0.0, 0.0, 0.0), voxel), np.uint8)
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "g.bin")
        subprocess.run([TOOL, "fogsphere", repr(radius), repr(voxel), out], check=True)
        ref = np.fromfile(out, np.uint8)
    assert mine.size == ref.size
    assert np.array_equal(mine, ref)
