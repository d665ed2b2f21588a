"""Writes tests/golden/nanovdb_*.npz: NanoVDB buffers and voxel lookups produced by the reference's own NanoVDB
32.3.3 GridBuilder and PNanoVDB accessor (oracle/_ref/nvdb_ref, built by `make -C oracle ref` from the reference
tree). Each fixture holds the dense input, the reference grid bytes, integer-coordinate lookups and nearest
world-space lookups (with the voxel each one selected). The nanovdb_fp16_*.npz fixtures hold grids of
GridBuilder<float, Fp16> (16-bit quantised leaves) and their integer-coordinate lookups.

The fog_sphere_default.npz fixture holds the reference module's default scene grid (createFogVolumeSphere) and its
lookups.

nvdb_file_block20.npz holds the block20 grid as a .nvdb file written by the reference's io::writeGrid.

Run from the repository root: python tests/golden/make_nanovdb_golden.py [fog_sphere | nvdb_file]
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, REPO)
from oracle.oracle import fbm_cloud  # noqa: E402

TOOL = os.path.join(REPO, "oracle", "_ref", "nvdb_ref")


def case_inputs():
    d = fbm_cloud(40, 36, 28, seed=3)
    fbm = ((d - d.min()) / (d.max() - d.min())).astype(np.float32)
    block = np.zeros((20, 20, 20), np.float32)
    block[1:-1, 1:-1, 1:-1] = 1.0  # createCloudBlock(8..) style constant block with an empty boundary layer
    wide = np.ones((12, 10, 136), np.float32)  # spans two lower nodes along x, first one pruned to a tile
    return {
        "fbm40": (fbm, 0.0125, (-0.25, -0.225, -0.175)),
        "block20": (block, 0.5 / 20, (-0.25, -0.25, -0.25)),
        "wide136": (wide, 0.5 / 136, (-0.25, -0.25 * 10 / 136, -0.25 * 12 / 136)),
    }


def run(*args):
    subprocess.run([TOOL, *map(str, args)], check=True)


def main():
    if not os.path.exists(TOOL):
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    rng = np.random.default_rng(2024)
    out_dir = os.path.dirname(os.path.abspath(__file__))
    for name, (dense, dx, origin) in case_inputs().items():
        nz, ny, nx = dense.shape
        with tempfile.TemporaryDirectory() as tmp:
            f = lambda n: os.path.join(tmp, n)  # noqa: E731
            dense.tofile(f("d.f32"))
            run("build", f("d.f32"), nx, ny, nz, repr(dx), *[repr(v) for v in origin], f("g.bin"))
            grid = np.fromfile(f("g.bin"), np.uint8)
            ijk = np.concatenate([
                rng.integers(-3, [nx + 3, ny + 3, nz + 3], size=(2000, 3)),
                np.array([[0, 0, 0], [nx - 1, ny - 1, nz - 1], [-1, 0, 0], [nx, ny, nz], [7, 7, 7], [8, 8, 8]]),
            ]).astype(np.int32)
            ijk.tofile(f("ijk.i32"))
            run("ijk", f("g.bin"), f("ijk.i32"), f("v.f32"))
            vals = np.fromfile(f("v.f32"), np.float32)
            lo = np.array(origin, np.float64)
            hi = lo + dx * np.array([nx, ny, nz], np.float64)
            xyz = rng.uniform(lo - 2 * dx, hi + 2 * dx, size=(2000, 3)).astype(np.float32)
            xyz.tofile(f("xyz.f32"))
            run("world", f("g.bin"), f("xyz.f32"), f("w.f32"), f("wijk.i32"))
            wvals = np.fromfile(f("w.f32"), np.float32)
            wijk = np.fromfile(f("wijk.i32"), np.int32).reshape(-1, 3)
        np.savez_compressed(os.path.join(out_dir, f"nanovdb_{name}.npz"), dense=dense, dx=np.float64(dx),
                            origin=np.array(origin, np.float64), grid=grid, ijk=ijk, values=vals, xyz=xyz,
                            world_values=wvals, world_ijk=wijk)
        print(name, "grid bytes", grid.size, "active lookups", int((vals != 0).sum()))
    rng = np.random.default_rng(2025)
    for name, (dense, dx, origin) in case_inputs().items():
        if name == "block20":
            continue  # one constant leaf value: pruned to tiles, nothing quantised
        nz, ny, nx = dense.shape
        with tempfile.TemporaryDirectory() as tmp:
            f = lambda n: os.path.join(tmp, n)  # noqa: E731
            dense.tofile(f("d.f32"))
            run("build16", f("d.f32"), nx, ny, nz, repr(dx), *[repr(v) for v in origin], f("g.bin"))
            grid = np.fromfile(f("g.bin"), np.uint8)
            ijk = np.concatenate([
                rng.integers(-3, [nx + 3, ny + 3, nz + 3], size=(2000, 3)),
                np.array([[0, 0, 0], [nx - 1, ny - 1, nz - 1], [-1, 0, 0], [nx, ny, nz], [7, 7, 7], [8, 8, 8]]),
            ]).astype(np.int32)
            ijk.tofile(f("ijk.i32"))
            run("ijk16", f("g.bin"), f("ijk.i32"), f("v.f32"))
            vals = np.fromfile(f("v.f32"), np.float32)
        np.savez_compressed(os.path.join(out_dir, f"nanovdb_fp16_{name}.npz"), dense=dense, dx=np.float64(dx),
                            origin=np.array(origin, np.float64), grid=grid, ijk=ijk, values=vals)
        print("fp16", name, "grid bytes", grid.size, "active lookups", int((vals != 0).sum()))


def fog_sphere():
    """fog_sphere_default.npz: the reference module's default scene grid, nanovdb::createFogVolumeSphere<float>(0.25,
    0, 0.01) (Module.cpp:793-795), and reference lookups over its index box and a margin (all voxels on the three
    axes and the diagonal, plus random voxels)."""
    if not os.path.exists(TOOL):
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    rng = np.random.default_rng(2026)
    out_dir = os.path.dirname(os.path.abspath(__file__))
    with tempfile.TemporaryDirectory() as tmp:
        f = lambda n: os.path.join(tmp, n)  # noqa: E731
        run("fogsphere", "0.25", "0.01", f("g.bin"))
        grid = np.fromfile(f("g.bin"), np.uint8)
        line = np.arange(-31, 32)
        zero = np.zeros_like(line)
        ijk = np.concatenate([
            np.stack([line, zero, zero], 1), np.stack([zero, line, zero], 1), np.stack([zero, zero, line], 1),
            np.stack([line, line, line], 1), rng.integers(-31, 32, size=(30000, 3)),
        ]).astype(np.int32)
        ijk.tofile(f("ijk.i32"))
        run("ijk", f("g.bin"), f("ijk.i32"), f("v.f32"))
        vals = np.fromfile(f("v.f32"), np.float32)
    np.savez_compressed(os.path.join(out_dir, "fog_sphere_default.npz"), grid=grid, ijk=ijk, values=vals)
    print("fog sphere grid bytes", grid.size, "nonzero lookups", int((vals != 0).sum()))


def nvdb_file():
    """nvdb_file_block20.npz: the block20 fixture's grid written as a .nvdb file by the reference's
    nanovdb::io::writeGrid (uncompressed), the format CloudData::loadFromNvdbFile reads (CloudData.cpp:1360-1365)."""
    if not os.path.exists(TOOL):
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    out_dir = os.path.dirname(os.path.abspath(__file__))
    grid = np.load(os.path.join(out_dir, "nanovdb_block20.npz"))["grid"]
    with tempfile.TemporaryDirectory() as tmp:
        g, f = os.path.join(tmp, "g.bin"), os.path.join(tmp, "g.nvdb")
        grid.tofile(g)
        run("nvdbfile", g, f)
        file_bytes = np.fromfile(f, np.uint8)
    np.savez_compressed(os.path.join(out_dir, "nvdb_file_block20.npz"), nvdb=file_bytes)
    print(".nvdb file bytes", file_bytes.size)


if __name__ == "__main__":
    if sys.argv[1:] == ["fog_sphere"]:
        fog_sphere()
    elif sys.argv[1:] == ["nvdb_file"]:
        nvdb_file()
    else:
        main()
        fog_sphere()
        nvdb_file()
