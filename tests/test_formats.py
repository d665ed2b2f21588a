"""CPU checks of the volume file formats the loaders read (host code, no GPU needed).

.xyz (CloudData::loadFromXyzFile, CloudData.cpp:236-316): three u32 sizes, three f64 voxel sizes, then the floats with
z fastest. The loader transposes to x fastest and min-max normalises with the minimum seeded with 0 when
max - min > 1e-6. The expectation below is written independently with numpy (a transposed array and the closed-form
normalisation), so a swapped axis or a dropped seed of the minimum fails.
"""
import ctypes
import os
import struct

import numpy as np
import pytest


def _reader():
    from cloudrendering_amd import _lib

    return _lib


def _write_xyz(path, field_zyx, voxel=(1.0, 1.0, 1.0), trailing=b""):
    """Writes the dense field given as [z][y][x] in the file order (z fastest: index z + (y + x*Y)*Z)."""
    nz, ny, nx = field_zyx.shape
    file_order = np.ascontiguousarray(field_zyx.transpose(2, 1, 0))  # [x][y][z]: z fastest
    with open(path, "wb") as f:
        f.write(struct.pack("<3I3d", nx, ny, nz, *voxel))
        f.write(file_order.astype("<f4").tobytes())
        f.write(trailing)


def _read(path, capacity=None):
    L = _reader()
    dims = (ctypes.c_int * 3)()
    vs = (ctypes.c_double * 3)()
    L.call("crvpt_read_xyz_file", path.encode(), dims, vs, None, 0)
    n = dims[0] * dims[1] * dims[2]
    out = np.zeros(n if capacity is None else capacity, np.float32)
    L.call("crvpt_read_xyz_file", path.encode(), dims, vs, out.ctypes.data, out.size)
    return tuple(dims), tuple(vs), out[:n].reshape(dims[2], dims[1], dims[0])


@pytest.mark.parametrize("shape", [(5, 3, 7), (1, 1, 4), (9, 6, 2)])
def test_xyz_transpose_and_normalisation(tmp_path, shape):
    rng = np.random.default_rng(sum(shape))
    f = (rng.random(shape, dtype=np.float32) * 3.0 + 0.5).astype(np.float32)  # all positive: min seeded with 0
    p = str(tmp_path / "a.xyz")
    _write_xyz(p, f, voxel=(0.5, 0.25, 2.0))
    dims, vs, got = _read(p)
    assert dims == (shape[2], shape[1], shape[0])
    assert vs == (0.5, 0.25, 2.0)
    mx = np.float32(f.max())
    expect = (f - np.float32(0.0)) / (mx - np.float32(0.0))
    np.testing.assert_array_equal(got, expect.astype(np.float32))


def test_xyz_negative_values_and_constant_field(tmp_path):
    f = np.linspace(-2.0, 6.0, 4 * 5 * 6, dtype=np.float32).reshape(4, 5, 6)
    p = str(tmp_path / "n.xyz")
    _write_xyz(p, f)
    _, _, got = _read(p)
    mn, mx = np.float32(-2.0), np.float32(f.max())
    np.testing.assert_array_equal(got, ((f - mn) / (mx - mn)).astype(np.float32))
    # a constant field of 0 is left as it is (max - min <= 1e-6)
    z = np.zeros((3, 3, 3), np.float32)
    _write_xyz(p, z)
    _, _, got = _read(p)
    np.testing.assert_array_equal(got, z)


def test_xyz_trailing_bytes_are_ignored(tmp_path):
    f = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
    p = str(tmp_path / "t.xyz")
    _write_xyz(p, f, trailing=b"\x01\x02\x03")
    _, _, got = _read(p)
    np.testing.assert_array_equal(got, f / np.float32(23.0))


@pytest.mark.parametrize("header,payload", [
    ((0, 4, 4), 0),                       # empty grid
    ((4, 4, 4), 4 * 4 * 4 * 4 - 4),       # truncated data
    ((0x80000000, 2, 2), 64),             # beyond int dimensions
    ((0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF), 64),  # size product overflows 64 bits
])
def test_xyz_rejects_bad_headers(tmp_path, header, payload):
    p = str(tmp_path / "bad.xyz")
    with open(p, "wb") as fh:
        fh.write(struct.pack("<3I3d", *header, 1.0, 1.0, 1.0))
        fh.write(b"\0" * payload)
    L = _reader()
    dims = (ctypes.c_int * 3)()
    vs = (ctypes.c_double * 3)()
    with pytest.raises(L.CrvptError, match="xyz"):
        L.call("crvpt_read_xyz_file", p.encode(), dims, vs, None, 0)


def test_xyz_truncated_header(tmp_path):
    p = str(tmp_path / "h.xyz")
    with open(p, "wb") as fh:
        fh.write(b"\0" * 20)
    L = _reader()
    with pytest.raises(L.CrvptError, match="header"):
        L.call("crvpt_read_xyz_file", p.encode(), (ctypes.c_int * 3)(), (ctypes.c_double * 3)(), None, 0)


# .nvdb (CloudData::loadFromNvdbFile, CloudData.cpp:1360-1365: nanovdb::io::readGrid of the file's first grid). The
# fixture is a file the reference's own nanovdb::io::writeGrid wrote (tests/golden/make_nanovdb_golden.py).
GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _read_nvdb(path):
    L = _reader()
    size = ctypes.c_uint64()
    L.call("crvpt_read_nvdb_file", path.encode(), None, 0, ctypes.byref(size))
    out = np.zeros(size.value, np.uint8)
    L.call("crvpt_read_nvdb_file", path.encode(), out.ctypes.data, out.size, ctypes.byref(size))
    return out


def test_nvdb_file_yields_the_reference_grid(tmp_path):
    f = np.load(os.path.join(GOLDEN_DIR, "nvdb_file_block20.npz"))["nvdb"]
    grid = np.load(os.path.join(GOLDEN_DIR, "nanovdb_block20.npz"))["grid"]
    p = str(tmp_path / "b.nvdb")
    f.tofile(p)
    got = _read_nvdb(p)
    assert got.size == grid.size
    np.testing.assert_array_equal(got, grid)


@pytest.mark.parametrize("damage", ["truncated_grid", "truncated_meta", "codec", "magic"])
def test_nvdb_file_rejects_damaged_files(tmp_path, damage):
    f = np.load(os.path.join(GOLDEN_DIR, "nvdb_file_block20.npz"))["nvdb"].copy()
    if damage == "truncated_grid":
        f = f[:-100]
    elif damage == "truncated_meta":
        f = f[:100]
    elif damage == "codec":
        f[14] = 1  # ZIP
    else:
        f[0] ^= 0xFF
    p = str(tmp_path / "d.nvdb")
    f.tofile(p)
    with pytest.raises(RuntimeError):
        _read_nvdb(p)
