"""NumPy restatement of the reference's EAW and SVGF denoisers — TEST INFRASTRUCTURE ONLY.

Used by the tests as the checker of the gfx950 denoiser kernels; nothing in the product imports it.

  eaw()   Data/Shaders/Denoiser/EAWDenoise.glsl:161-314 (compute path) with the host loop of
          src/Denoiser/EAWDenoiser.cpp:273-341 (step widths 1, 2, 4, ...; colour weight min(exp(-1/phiColor), 1)
          independent of the colour distance; position and normal weights min(exp(-|d|^2/phi), 1)); taps outside the
          image clamp to the edge.
  Svgf    Data/Shaders/Denoiser/SVGF/SVGF.glsl (Compute-Reproject, Compute-Filter-Moments, Compute-ATrous) and
          svgf_common.glsl compute_weight, sequenced as SVGFDenoiser::denoise (src/Denoiser/SVGF.cpp:118-180): the
          colour history is the first a-trous iteration's output, then normal, depth and moments become history.
          The moments filter reads its input from a separate buffer (the reference reads and writes one image in
          place, a race whose outcome depends on the dispatch order).
All arithmetic is float32 per pixel, vectorised over the image.
"""
from __future__ import annotations

import numpy as np

F = np.float32


def _tap(img, y, x):
    """img[y, x] with coordinates clamped to the image (CLAMP_TO_EDGE)."""
    h, w = img.shape[:2]
    return img[np.clip(y, 0, h - 1), np.clip(x, 0, w - 1)]


def eaw(color, position=None, normal=None, iterations=0, phi_color=5.0, phi_position=0.1, phi_normal=0.1,
        use_color=True, use_position=True, use_normal=True):
    """color, position, normal: [H, W, 4] float32. Returns the filtered [H, W, 4] image."""
    out = np.asarray(color, F).copy()
    h, w = out.shape[:2]
    yy, xx = np.mgrid[0:h, 0:w]
    step = 1
    wc = F(min(np.exp(F(-1.0) / F(phi_color)), F(1.0)))
    for _ in range(iterations):
        c0 = out
        s = np.zeros_like(out)
        acc = np.zeros((h, w), F)
        for i in range(25):
            ox, oy = i % 5 - 2, i // 5 - 2
            k = F(np.exp(F(-float(ox * ox + oy * oy)) / F(2.0)))
            c = _tap(c0, yy + oy * step, xx + ox * step)
            wt = np.ones((h, w), F)
            if use_color:
                wt = wt * wc
            if position is not None and use_position:
                p0 = np.asarray(position, F)
                d = p0 - _tap(p0, yy + oy * step, xx + ox * step)
                wt = wt * np.minimum(np.exp(-(d * d).sum(-1, dtype=F) / F(phi_position)), F(1.0)).astype(F)
            if normal is not None and use_normal:
                n0 = np.asarray(normal, F)
                d = n0 - _tap(n0, yy + oy * step, xx + ox * step)
                wt = wt * np.minimum(np.exp(-(d * d).sum(-1, dtype=F) / F(phi_normal)), F(1.0)).astype(F)
            s = s + c * (wt * k)[..., None]
            acc = acc + wt * k
        out = (s / acc[..., None]).astype(F)
        step *= 2
    return out


def _compute_weight(cd, od, phi_d, cn, on, cc, oc, phi_c):
    """svgf_common.glsl compute_weight (sigma_n = 128). GLSL max() of a NaN and 0 returns 0 on the GPU (IEEE
    maxNum), which np.fmax reproduces: a zero colour scale with equal colours (0 / 0) gives no colour penalty."""
    wn = np.power(np.fmax(F(0.0), (cn * on).sum(-1, dtype=F)), F(128.0)).astype(F)
    with np.errstate(divide="ignore", invalid="ignore"):
        wz = np.where(phi_d == 0, F(0.0), np.abs(cd - od) / np.where(phi_d == 0, F(1.0), phi_d)).astype(F)
        wcol = (np.abs(cc - oc) * F(2.0) / phi_c).astype(F)
    return (np.exp(F(0.0) - np.fmax(wcol, F(0.0)) - np.fmax(wz, F(0.0))) * wn).astype(F)


def _fetch(img, y, x, fill=0.0):
    """texelFetch with zero outside the image."""
    h, w = img.shape[:2]
    inside = (y >= 0) & (y < h) & (x >= 0) & (x < w)
    v = img[np.clip(y, 0, h - 1), np.clip(x, 0, w - 1)]
    if v.ndim == 3:
        return np.where(inside[..., None], v, F(fill)), inside
    return np.where(inside, v, F(fill)), inside


class Svgf:
    """SVGF state across frames (history textures start at zero, SVGF.cpp:244-268)."""

    def __init__(self, h, w, iterations=5, allowed_z_dist=0.002, allowed_normal_dist=0.02):
        self.h, self.w, self.iterations = h, w, iterations
        self.az, self.an = F(allowed_z_dist), F(allowed_normal_dist)
        self.color_history = np.zeros((h, w, 4), F)
        self.moments_history = np.zeros((h, w, 4), F)
        self.prev_normal = np.zeros((h, w, 4), F)
        self.prev_depth = np.zeros((h, w), F)

    def _valid(self, y, x, z, zp, n, np_):
        inside = (x >= 1) & (y >= 1) & (x <= self.w - 1) & (y <= self.h - 1)
        dn = np.sqrt(((np_ - n) ** 2).sum(-1, dtype=F)).astype(F)
        return inside & ~(np.abs(zp - z) > self.az) & ~(dn > self.an)

    def reproject(self, noisy, motion, normal, depth):
        h, w = self.h, self.w
        yy, xx = np.mgrid[0:h, 0:w]
        mx, my = motion[..., 0], motion[..., 1]
        # ivec2(vec2(0.5) + i_pos - motion): conversion truncates toward zero
        px = np.trunc(F(0.5) + xx.astype(F) - mx).astype(np.int64)
        py = np.trunc(F(0.5) + yy.astype(F) - my).astype(np.int64)
        load_ok = (px >= 0) & (py >= 0) & (px < w) & (py < h)
        mh, _ = _fetch(self.moments_history, py, px)
        prev_m = mh[..., :2].copy()
        hist_len = mh[..., 2].copy()
        color = noisy[..., :3].astype(F)
        clf = self.color_history[..., :3].copy()  # texelFetch(color_history, i_pos)
        success = load_ok.copy()
        fx = F(0.01) + xx.astype(F) - mx
        fy = F(0.01) + yy.astype(F) - my
        ix = np.trunc(fx).astype(np.int64)
        iy = np.trunc(fy).astype(np.int64)
        z = depth.astype(F)
        n = normal[..., :3].astype(F)
        # try_2x2_tap: offsets (0,0), (0,1), (1,0), (1,1) with the weights of (0,0), (1,0), (0,1), (1,1)
        offs = [(0, 0), (0, 1), (1, 0), (1, 1)]
        fxr, fyr = fx - np.floor(fx), fy - np.floor(fy)
        wts = [(1 - fxr) * (1 - fyr), fxr * (1 - fyr), (1 - fxr) * fyr, fxr * fyr]
        valids = []
        for ox, oy in offs:
            on, _ = _fetch(self.prev_normal, iy + oy, ix + ox)
            od, _ = _fetch(self.prev_depth, iy + oy, ix + ox)
            valids.append(self._valid(iy + oy, ix + ox, z, od, n, on[..., :3]))
        any_valid = valids[0] | valids[1] | valids[2] | valids[3]
        cb = np.zeros((h, w, 3), F)
        mb = np.zeros((h, w, 2), F)
        sw = np.zeros((h, w), F)
        for (ox, oy), wt, v in zip(offs, wts, valids):
            oc, _ = _fetch(self.color_history, iy + oy, ix + ox)
            om, _ = _fetch(self.moments_history, iy + oy, ix + ox)
            wt = np.where(v, wt, F(0.0)).astype(F)
            cb = cb + wt[..., None] * oc[..., :3]
            mb = mb + wt[..., None] * om[..., :2]
            sw = sw + wt
        ok2 = success & any_valid & (sw >= F(0.001))
        sws = np.where(ok2, sw, F(1.0))
        clf = np.where(ok2[..., None], cb / sws[..., None], clf)
        prev_m = np.where(ok2[..., None], mb / sws[..., None], prev_m)
        # try_3x3_bilat where the 2x2 tap failed
        need3 = success & ~ok2
        fc = np.zeros((h, w, 3), F)
        fm = np.zeros((h, w, 2), F)
        nv = np.zeros((h, w), F)
        for oy in (-1, 0, 1):
            for ox in (-1, 0, 1):
                qx, qy = ix + ox, iy + oy
                inside = (qx >= 1) & (qy >= 1) & (qx < w) & (qy < h)
                on, _ = _fetch(self.prev_normal, qy, qx)
                od, _ = _fetch(self.prev_depth, qy, qx)
                v = inside & self._valid(qy, qx, z, od, n, on[..., :3])
                oc, _ = _fetch(self.color_history, qy, qx)
                om, _ = _fetch(self.moments_history, qy, qx)
                fc = fc + np.where(v[..., None], oc[..., :3], F(0.0))
                fm = fm + np.where(v[..., None], om[..., :2], F(0.0))
                nv = nv + v.astype(F)
        ok3 = need3 & (nv > 0)
        nvs = np.where(ok3, nv, F(1.0))
        clf = np.where(ok3[..., None], fc / nvs[..., None], clf)
        prev_m = np.where(ok3[..., None], fm / nvs[..., None], prev_m)
        success = ok2 | ok3
        prev_m = np.where(load_ok[..., None] | success[..., None], prev_m, F(0.0))
        hl = np.minimum(np.where(success, hist_len + F(1.0), F(1.0)), F(32.0)).astype(F)
        ac = np.where(success, np.maximum(F(0.01), F(1.0) / hl), F(1.0)).astype(F)
        am = np.where(success, np.maximum(F(0.2), F(1.0) / hl), F(1.0)).astype(F)
        m0 = color[..., 0]
        m1 = m0 * m0
        mr = prev_m[..., 0] * (1 - am) + m0 * am
        mg = prev_m[..., 1] * (1 - am) + m1 * am
        var = np.maximum(F(0.0), mg - mr * mr)
        accum_moments = np.stack([mr, mg, hl, np.zeros_like(hl)], -1).astype(F)
        temp = np.concatenate([clf * (1 - ac[..., None]) + color * ac[..., None], var[..., None]], -1).astype(F)
        return accum_moments, temp

    def filter_moments(self, temp, accum_moments, normal, depth, fwidth):
        h, w = self.h, self.w
        yy, xx = np.mgrid[0:h, 0:w]
        hl = accum_moments[..., 2]
        cd = depth.astype(F)
        cf = np.abs(fwidth).astype(F) + F(0.0001)
        cn = normal[..., :3].astype(F)
        cc = temp[..., 0]
        sw = np.zeros((h, w), F)
        sc = np.zeros((h, w, 3), F)
        sm = np.zeros((h, w, 2), F)
        for oy in range(-3, 4):
            for ox in range(-3, 4):
                qy, qx = yy + oy, xx + ox
                inside = (qx >= 0) & (qy >= 0) & (qx < w) & (qy < h)
                oc, _ = _fetch(temp, qy, qx)
                om, _ = _fetch(accum_moments, qy, qx)
                od, _ = _fetch(depth.astype(F), qy, qx)
                on, _ = _fetch(normal.astype(F), qy, qx)
                wt = _compute_weight(cd, od, cf, cn, on[..., :3], cc, oc[..., 0], F(10.0))
                wt = np.where(inside, wt, F(0.0)).astype(F)
                sw = sw + wt
                sc = sc + wt[..., None] * oc[..., :3]
                sm = sm + wt[..., None] * om[..., :2]
        sw = np.maximum(sw, F(1e-6))
        sc = sc / sw[..., None]
        sm = sm / sw[..., None]
        var = (sm[..., 1] - sm[..., 0] * sm[..., 0]) * (F(4.0) / np.where(hl > 0, hl, F(1.0)))
        out = np.concatenate([sc, var[..., None]], -1).astype(F)
        return np.where((hl >= 4)[..., None], temp, out).astype(F)

    def atrous(self, color, normal, depth, fwidth, iteration):
        h, w = self.h, self.w
        yy, xx = np.mgrid[0:h, 0:w]
        step = 1 << iteration
        kern = {(0, 0): F(1 / 4), (0, 1): F(1 / 8), (1, 0): F(1 / 8), (1, 1): F(1 / 16)}
        fv = np.zeros((h, w), F)
        for oy in (-1, 0, 1):
            for ox in (-1, 0, 1):
                c, _ = _fetch(color, yy + oy, xx + ox)
                fv = fv + c[..., 3] * kern[(abs(ox), abs(oy))]
        phi_c = np.sqrt(np.maximum(F(0.0), F(1e-10) + fv)).astype(F)
        kv = [F(1.0), F(2.0 / 3.0), F(1.0 / 6.0)]
        acc = np.full((h, w), kv[0] * kv[0], F)
        s = color * acc[..., None]
        cn = normal[..., :3].astype(F)
        cz = depth.astype(F)
        cf = fwidth.astype(F)
        for y in range(-2, 3):
            for x in range(-2, 3):
                if x == 0 and y == 0:
                    continue
                qy, qx = yy + y * step, xx + x * step
                inside = (qx >= 0) & (qy >= 0) & (qx < w) & (qy < h)
                k = kv[abs(x)] * kv[abs(y)]
                oc, _ = _fetch(color, qy, qx)
                on, _ = _fetch(normal.astype(F), qy, qx)
                oz, _ = _fetch(depth.astype(F), qy, qx)
                phi_d = (np.abs(cf * F(np.sqrt(float(x * x + y * y))) * F(step)) + F(0.0001)).astype(F)
                wt = _compute_weight(cz, oz, phi_d, cn, on[..., :3], color[..., 0], oc[..., 0], phi_c) * k
                wt = np.where(inside, wt, F(0.0)).astype(F)
                s = s + np.stack([wt, wt, wt, wt * wt], -1) * oc
                acc = acc + wt
        return (s / np.stack([acc, acc, acc, acc * acc], -1)).astype(F)

    def denoise(self, noisy, motion, normal, depth, fwidth):
        """One frame: noisy [H, W, 4], motion [H, W, 2], normal [H, W, 4], depth [H, W], fwidth [H, W]."""
        accum_moments, temp = self.reproject(noisy, motion, normal, depth)
        temp = self.filter_moments(temp, accum_moments, normal, depth, fwidth)
        if self.iterations < 1:
            out = temp
            self.color_history = temp.copy()
        else:
            cur = temp
            for i in range(self.iterations):
                cur = self.atrous(cur, normal, depth, fwidth, i)
                if i == 0:
                    self.color_history = cur.copy()
            out = cur
        self.prev_normal = np.asarray(normal, F).copy()
        self.prev_depth = np.asarray(depth, F).copy()
        self.moments_history = accum_moments.copy()
        return out
