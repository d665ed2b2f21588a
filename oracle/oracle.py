"""ctypes front end of the CPU oracle (liboracle.so, built from vpt_oracle.cpp by oracle/Makefile).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, always as
the checker or the timed CPU baseline, never as a product path.

Also holds a NumPy restatement of the build's seeded fBm cloud generator (the GPU generator in render.hip) so
CPU-only tests can make the same volumes.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

MODES = {"delta": 0, "delta_spectral": 1, "ratio": 2, "decomposition": 3, "residual_ratio": 4, "next_event": 5,
         "next_event_spectral": 6}
INTERP = {"nearest": 0, "stochastic": 1, "trilinear": 2}
ENVS = {"default": 0, "black": 1, "single_color": 2, "image": 3}
T_ESTIMATORS = {"delta": 0, "ratio": 1, "residual_ratio": 2, "decomposition": 3}


class OracleLight(ctypes.Structure):
    _fields_ = [
        ("light_type", ctypes.c_uint32), ("light_space", ctypes.c_uint32),
        ("spot_total_width", ctypes.c_float), ("spot_falloff_start", ctypes.c_float),
        ("color", ctypes.c_float * 3), ("intensity", ctypes.c_float),
        ("position", ctypes.c_float * 3), ("use_distance", ctypes.c_uint32),
        ("spot_direction", ctypes.c_float * 3), ("padding", ctypes.c_float),
    ]


class OracleScene(ctypes.Structure):
    _fields_ = [
        ("density", ctypes.POINTER(ctypes.c_float)),
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("voxel_value_min", ctypes.c_float), ("voxel_value_max", ctypes.c_float),
        ("box_min", ctypes.c_float * 3), ("box_max", ctypes.c_float * 3),
        ("interpolation", ctypes.c_int),
        ("extinction", ctypes.c_float * 3), ("scattering_albedo", ctypes.c_float * 3), ("phase_g", ctypes.c_float),
        ("env_kind", ctypes.c_int),
        ("sun_direction", ctypes.c_float * 3), ("sun_intensity", ctypes.c_float * 3),
        ("env_image", ctypes.POINTER(ctypes.c_float)), ("env_w", ctypes.c_int), ("env_h", ctypes.c_int),
        ("env_intensity_factor", ctypes.c_float * 3), ("is_env_map_black", ctypes.c_int),
        ("use_linear_rgb", ctypes.c_int),
        ("num_lights", ctypes.c_int), ("lights", ctypes.POINTER(OracleLight)),
        ("inverse_view_proj", ctypes.c_float * 16), ("inverse_view", ctypes.c_float * 16),
        ("mode", ctypes.c_int), ("custom_seed_offset", ctypes.c_uint32), ("super_voxel_size", ctypes.c_int),
        ("max_bounces", ctypes.c_int),
    ]


_lib = None


def build() -> str:
    """Compiles liboracle.so with the committed Makefile (g++, OpenMP)."""
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    lib.oracle_render.argtypes = [P(OracleScene), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    lib.oracle_render.restype = ctypes.c_int
    lib.oracle_render_sums.argtypes = [P(OracleScene), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    lib.oracle_render_sums.restype = ctypes.c_int
    lib.oracle_transmittance.argtypes = [P(OracleScene), ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    lib.oracle_transmittance.restype = ctypes.c_int
    lib.oracle_optical_depth_batch.argtypes = [P(OracleScene), ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    lib.oracle_optical_depth_batch.restype = ctypes.c_int
    lib.oracle_supervoxel_tables.argtypes = [P(OracleScene), ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p]
    lib.oracle_supervoxel_tables.restype = ctypes.c_int
    lib.oracle_render_features.argtypes = [P(OracleScene), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p, ctypes.c_int]
    lib.oracle_render_features.restype = ctypes.c_int
    lib.oracle_hybrid_taus.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p]
    lib.oracle_trace_rays.argtypes = [P(OracleScene), ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    lib.oracle_trace_rays.restype = ctypes.c_int
    lib.oracle_sky.argtypes = [P(OracleScene), ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    lib.oracle_sky.restype = ctypes.c_int
    lib.oracle_light_radiance.argtypes = [P(OracleScene), ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_void_p]
    lib.oracle_light_radiance.restype = ctypes.c_int
    lib.oracle_camera_matrices.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_double] * 4 + [ctypes.c_void_p] * 3
    _lib = lib
    return lib


def camera_matrices(pos=(0.0, 0.0, 0.8), target=None, up=(0.0, 1.0, 0.0), fovy=2.0 * math.atan(0.5), aspect=1.0,
                    znear=1.0 / 512.0, zfar=80.0):
    """(inverse view-projection, inverse view, view-projection) as float32[16] column-major. The default camera is
    the reference module's (VolumetricPathTracingModuleRenderer.cpp:124-131): at (0, 0, 0.8) looking down -z."""
    lib = load()
    if target is None:
        target = (pos[0], pos[1], pos[2] - 1.0)
    p = (ctypes.c_double * 3)(*pos)
    t = (ctypes.c_double * 3)(*target)
    u = (ctypes.c_double * 3)(*up)
    ivp = np.zeros(16, np.float32)
    iv = np.zeros(16, np.float32)
    vp = np.zeros(16, np.float32)
    lib.oracle_camera_matrices(p, t, u, fovy, aspect, znear, zfar, ivp.ctypes.data, iv.ctypes.data, vp.ctypes.data)
    return ivp, iv, vp


def box_for_grid(shape_zyx, max_grid_extent=0.25):
    """CloudData::computeGridBounds (CloudData.cpp:63-78) with unit voxel sizes."""
    nz, ny, nx = shape_zyx
    s = np.array([nx, ny, nz], np.float32)
    bmax = s * np.float32(max_grid_extent) / np.float32(s.max())
    return -bmax, bmax


# Feature maps and their channel counts in the reference's image formats (VolumetricPathTracingPass.cpp:295-445).
FEATURE_MAPS = {"result": 4, "first_x": 4, "first_w": 4, "normal": 4, "cloud_only": 4, "background": 4, "depth": 2,
                "depth_far": 1, "density": 2, "reproj_uv": 2, "albedo": 4, "flow": 2, "flow_reverse": 2,
                "depth_nabla": 2, "depth_fwidth": 1}


class OracleFeatureMaps(ctypes.Structure):
    _fields_ = [(name, ctypes.POINTER(ctypes.c_float)) for name in FEATURE_MAPS]


class Scene:
    """Holds an OracleScene plus the numpy buffers it points at."""

    def __init__(self, density_zyx, mode="delta", interpolation="stochastic", extinction=1024.0, albedo=0.9,
                 g=0.5, env="default", env_image=None, env_factor=(1.0, 1.0, 1.0), lights=(),
                 camera=None, width=64, height=64, seed_offset=0, max_grid_extent=0.25, box=None,
                 sun_direction=(0.5826, 0.7660, 0.2717), sun_intensity=2.6,
                 sun_color=(1.0, 0.961538462, 0.884615385), super_voxel_size=8, use_linear_rgb=True,
                 voxel_range=None, max_bounces=0):
        self.density = np.ascontiguousarray(np.asarray(density_zyx, np.float32))
        nz, ny, nx = self.density.shape
        sc = OracleScene()
        sc.density = self.density.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        sc.nx, sc.ny, sc.nz = nx, ny, nz
        if voxel_range is None:
            mn, mx = float(self.density.min()), float(self.density.max())
            if mn == mx:
                mn = 0.0
            voxel_range = (mn, mx)
        sc.voxel_value_min, sc.voxel_value_max = voxel_range
        bmin, bmax = box if box is not None else box_for_grid(self.density.shape, max_grid_extent)
        sc.box_min[:] = [float(v) for v in bmin]
        sc.box_max[:] = [float(v) for v in bmax]
        sc.interpolation = INTERP[interpolation] if isinstance(interpolation, str) else int(interpolation)
        ext = extinction if np.ndim(extinction) else (extinction, extinction, extinction)
        alb = albedo if np.ndim(albedo) else (albedo, albedo, albedo)
        sc.extinction[:] = [float(v) for v in ext]
        sc.scattering_albedo[:] = [float(v) for v in alb]
        sc.phase_g = g
        sd = np.array(sun_direction, np.float64)
        sd /= np.linalg.norm(sd)
        sc.sun_direction[:] = [float(v) for v in sd]
        sc.sun_intensity[:] = [float(sun_intensity * c) for c in sun_color]
        self.env_image = None
        if env_image is not None:
            img = np.asarray(env_image, np.float32)  # [C, H, W]
            c, h, w = img.shape
            rgba = np.ones((h, w, 4), np.float32)
            rgba[..., :3] = np.moveaxis(img[:3], 0, -1)
            if c == 4:
                rgba[..., 3] = img[3]
            self.env_image = np.ascontiguousarray(rgba)
            sc.env_image = self.env_image.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            sc.env_w, sc.env_h = w, h
            sc.env_kind = ENVS["image"]
            black = not np.any(img[:3] != 0)
            if all(f < 1e-5 for f in env_factor):
                black = True
            sc.is_env_map_black = int(black)
        else:
            sc.env_kind = ENVS[env]
            sc.is_env_map_black = int(env == "black")
        sc.env_intensity_factor[:] = [float(v) for v in env_factor]
        sc.use_linear_rgb = int(use_linear_rgb)
        self.lights = (OracleLight * max(1, len(lights)))()
        for i, L in enumerate(lights):
            self.lights[i] = L
        sc.num_lights = len(lights)
        sc.lights = ctypes.cast(self.lights, ctypes.POINTER(OracleLight))
        if camera is None:
            camera = camera_matrices(aspect=width / height)
        ivp, iv = camera[0], camera[1]
        self.view_proj = camera[2] if len(camera) > 2 else None
        sc.inverse_view_proj[:] = [float(v) for v in ivp]
        sc.inverse_view[:] = [float(v) for v in iv]
        sc.mode = MODES[mode] if isinstance(mode, str) else int(mode)
        sc.custom_seed_offset = seed_offset & 0xFFFFFFFF
        sc.super_voxel_size = super_voxel_size
        sc.max_bounces = int(max_bounces)
        self.sc = sc
        self.width, self.height = width, height

    def render(self, frames, frame_start=0, acc=None, threads=0):
        """Running-mean image [H, W, 3] after frames [frame_start, frame_start + frames), plus per-pixel sum of
        squared path radiance (for variance estimates)."""
        lib = load()
        h, w = self.height, self.width
        if acc is None:
            acc = np.zeros((h, w, 3), np.float32)
        sq = np.zeros((h, w, 3), np.float64)
        rc = lib.oracle_render(ctypes.byref(self.sc), w, h, frame_start, frames, acc.ctypes.data, sq.ctypes.data,
                               threads)
        assert rc == 0
        return acc, sq

    def render_mean(self, frames, frame_start=0, threads=0):
        """Exact FP64 sample mean [H, W, 3] of frames [frame_start, frame_start + frames) (the samples of render),
        plus the per-pixel sum of squared path radiance. Converged comparisons use this: the reference's FP32
        running mean drifts by up to a few 1e-4 relative over 10^4-10^5 frames."""
        lib = load()
        h, w = self.height, self.width
        s1 = np.zeros((h, w, 3), np.float64)
        sq = np.zeros((h, w, 3), np.float64)
        rc = lib.oracle_render_sums(ctypes.byref(self.sc), w, h, frame_start, frames, s1.ctypes.data,
                                    sq.ctypes.data, threads)
        assert rc == 0
        return s1 / max(frames, 1), sq

    def render_features(self, frames, names=FEATURE_MAPS, frame_start=0, state=None, prev_view_proj=None,
                        far_distance=80.0, threads=0):
        """Feature maps of Clouds.glsl after frames [frame_start, frame_start + frames): a dict name -> [H, W, C]
        float32 in the reference image format's channel count; `state` (a previous result) continues it."""
        lib = load()
        h, w = self.height, self.width
        maps = OracleFeatureMaps()
        out = {}
        for name in names:
            ch = FEATURE_MAPS[name]
            a = state[name] if state is not None else np.zeros((h, w, ch), np.float32)
            a = np.ascontiguousarray(a, np.float32)
            out[name] = a
            setattr(maps, name, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        if prev_view_proj is None:
            prev_view_proj = np.asarray(self.view_proj, np.float32)
        pvp = np.ascontiguousarray(np.asarray(prev_view_proj, np.float32).reshape(16))
        rc = lib.oracle_render_features(ctypes.byref(self.sc), w, h, frame_start, frames, pvp.ctypes.data,
                                        float(far_distance), ctypes.byref(maps), threads)
        assert rc == 0
        return out

    def transmittance(self, estimator, origins, directions, t_max=None, n_samples=1, seed=1, threads=0):
        lib = load()
        o = np.ascontiguousarray(np.asarray(origins, np.float32).reshape(-1, 3))
        d = np.ascontiguousarray(np.asarray(directions, np.float32).reshape(-1, 3))
        n = o.shape[0]
        tm = None if t_max is None else np.ascontiguousarray(np.asarray(t_max, np.float32).reshape(n))
        mean = np.zeros(n, np.float64)
        m2 = np.zeros(n, np.float64)
        est = T_ESTIMATORS[estimator] if isinstance(estimator, str) else int(estimator)
        rc = lib.oracle_transmittance(ctypes.byref(self.sc), est, n, o.ctypes.data, d.ctypes.data,
                                      None if tm is None else tm.ctypes.data, n_samples, seed, mean.ctypes.data,
                                      m2.ctypes.data, threads)
        assert rc == 0
        return mean, m2

    def trace_rays(self, origins, directions, seed=1, rrt_max_iterations=-1, threads=0):
        """One path of the scene's tracker per ray from (origin, direction): radiance [n, 3] float32 and, for
        residual ratio tracking, the in-scattering iterations made [n] (see oracle_trace_rays)."""
        lib = load()
        o = np.ascontiguousarray(np.asarray(origins, np.float32).reshape(-1, 3))
        d = np.ascontiguousarray(np.asarray(directions, np.float32).reshape(-1, 3))
        n = o.shape[0]
        rgb = np.zeros((n, 3), np.float32)
        it = np.zeros(n, np.int32)
        rc = lib.oracle_trace_rays(ctypes.byref(self.sc), n, o.ctypes.data, d.ctypes.data, seed & 0xFFFFFFFF,
                                   int(rrt_max_iterations), rgb.ctypes.data, it.ctypes.data, threads)
        assert rc == 0
        return rgb, it

    def sky(self, directions):
        """sampleSkybox + sampleLight for unit directions [n, 3] -> [n, 3] float32."""
        lib = load()
        d = np.ascontiguousarray(np.asarray(directions, np.float32).reshape(-1, 3))
        out = np.zeros_like(d)
        assert lib.oracle_sky(ctypes.byref(self.sc), d.shape[0], d.ctypes.data, out.ctypes.data) == 0
        return out

    def light_radiance(self, light_idx, positions):
        """sampleHeadlight: radiance of light light_idx towards positions [n, 3] -> [n, 3] float32."""
        lib = load()
        p = np.ascontiguousarray(np.asarray(positions, np.float32).reshape(-1, 3))
        out = np.zeros_like(p)
        assert lib.oracle_light_radiance(ctypes.byref(self.sc), int(light_idx), p.shape[0], p.ctypes.data,
                                         out.ctypes.data) == 0
        return out

    def optical_depth(self, origins, directions, t_max=None, threads=0):
        lib = load()
        o = np.ascontiguousarray(np.asarray(origins, np.float32).reshape(-1, 3))
        d = np.ascontiguousarray(np.asarray(directions, np.float32).reshape(-1, 3))
        n = o.shape[0]
        tm = None if t_max is None else np.ascontiguousarray(np.asarray(t_max, np.float32).reshape(n))
        tau = np.zeros(n, np.float64)
        lib.oracle_optical_depth_batch(ctypes.byref(self.sc), n, o.ctypes.data, d.ctypes.data,
                                       None if tm is None else tm.ctypes.data, tau.ctypes.data, threads)
        return tau

    def supervoxel_tables(self, extinction_x=None):
        lib = load()
        ext = float(self.sc.extinction[0] if extinction_x is None else extinction_x)
        nz, ny, nx = self.density.shape
        g = np.zeros(3, np.int32)
        n = (nx + 7) * (ny + 7) * (nz + 7)  # upper bound of the table size for any supervoxel edge >= 1
        rrt = np.zeros(2 * n, np.float32)
        dec = np.zeros(2 * n, np.float32)
        occ = np.zeros(n, np.uint8)
        sv = lib.oracle_supervoxel_tables(ctypes.byref(self.sc), ext, g.ctypes.data, rrt.ctypes.data,
                                          dec.ctypes.data, occ.ctypes.data)
        m = int(g.prod())
        return sv, tuple(int(v) for v in g), rrt[:2 * m].reshape(-1, 2), dec[:2 * m].reshape(-1, 2), occ[:m]


def hybrid_taus(seed, n):
    lib = load()
    out = np.zeros(n, np.float32)
    lib.oracle_hybrid_taus(seed & 0xFFFFFFFF, n, out.ctypes.data)
    return out


def make_light(light_type=2, light_space=0, color=(1.0, 1.0, 1.0), intensity=1.0, position=(0.0, 0.0, 0.0),
               use_distance=1, spot_total_width=0.0981747704, spot_falloff_start=0.0245436926,
               spot_direction=(0.0, 0.0, -1.0)):
    L = OracleLight()
    L.light_type, L.light_space = light_type, light_space
    L.spot_total_width, L.spot_falloff_start = spot_total_width, spot_falloff_start
    L.color[:] = list(color)
    L.intensity = intensity
    L.position[:] = list(position)
    L.use_distance = use_distance
    L.spot_direction[:] = list(spot_direction)
    return L


# ---------------------------------------------------------------------------------------------------------------
# NumPy restatement of the build's fBm generator (k_fbm in render.hip); every operation is one IEEE float32 op.
# ---------------------------------------------------------------------------------------------------------------

def _lowbias32(x):
    x = x.astype(np.uint32)
    x ^= x >> np.uint32(16)
    x = (x * np.uint32(0x7FEB352D)).astype(np.uint32)
    x ^= x >> np.uint32(15)
    x = (x * np.uint32(0x846CA68B)).astype(np.uint32)
    x ^= x >> np.uint32(16)
    return x


def _lattice(ix, iy, iz, salt):
    h = _lowbias32((ix.astype(np.uint32) * np.uint32(73856093)) ^ (iy.astype(np.uint32) * np.uint32(19349663))
                   ^ (iz.astype(np.uint32) * np.uint32(83492791)) ^ np.uint32(salt))
    return (h >> np.uint32(8)).astype(np.float32) * np.float32(1.0 / 16777216.0)


def _fade(t):
    f = np.float32
    return t * t * t * (t * (t * f(6.0) - f(15.0)) + f(10.0))


def _value_noise(px, py, pz, salt):
    fx, fy, fz = np.floor(px), np.floor(py), np.floor(pz)
    ix, iy, iz = fx.astype(np.int32), fy.astype(np.int32), fz.astype(np.int32)
    u, v, w = _fade(px - fx), _fade(py - fy), _fade(pz - fz)
    c = [_lattice(ix + (k & 1), iy + ((k >> 1) & 1), iz + ((k >> 2) & 1), salt) for k in range(8)]
    x00 = c[0] + (c[1] - c[0]) * u
    x10 = c[2] + (c[3] - c[2]) * u
    x01 = c[4] + (c[5] - c[4]) * u
    x11 = c[6] + (c[7] - c[6]) * u
    y0 = x00 + (x10 - x00) * v
    y1 = x01 + (x11 - x01) * v
    return y0 + (y1 - y0) * w


def fbm_cloud(nx, ny, nz, seed=0, octaves=5, lacunarity=2.0, gain=0.5, base_frequency=4.0, edge0=0.45, edge1=0.6,
              height_falloff=False):
    """[nz, ny, nx] float32 volume, identical operation by operation to the GPU generator."""
    f = np.float32
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    px = (x.astype(f) + f(0.5)) / f(nx)
    py = (y.astype(f) + f(0.5)) / f(ny)
    pz = (z.astype(f) + f(0.5)) / f(nz)
    s = np.zeros_like(px)
    norm, amp, freq = f(0.0), f(1.0), f(base_frequency)
    for o in range(octaves):
        salt = (np.uint64(seed) * np.uint64(2654435761) + np.uint64(o) * np.uint64(0x9E3779B9)) & np.uint64(0xFFFFFFFF)
        s = s + amp * _value_noise(px * freq, py * freq, pz * freq, int(salt))
        norm = f(norm + amp)
        amp = f(amp * f(gain))
        freq = f(freq * f(lacunarity))
    v = s / norm
    if height_falloff:
        h = py
        base = np.clip((h - f(0.05)) / f(0.25), f(0.0), f(1.0))
        top = np.clip((f(0.95) - h) / f(0.1), f(0.0), f(1.0))
        spread = f(1.0) + f(0.5) * np.clip((h - f(0.6)) / f(0.3), f(0.0), f(1.0))
        v = v * base * top * spread
    t = np.clip((v - f(edge0)) / f(f(edge1) - f(edge0)), f(0.0), f(1.0))
    return (t * t * (f(3.0) - f(2.0) * t)).astype(np.float32)
