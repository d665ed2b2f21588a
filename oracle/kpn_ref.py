"""NumPy restatement of the kernel-prediction apply kernels — TEST INFRASTRUCTURE ONLY.

  per_pixel_kernel  modules/kpn_module/src/Kernel.cu:38-131: out[n, c, y, x] = sum_k w[n, k, y, x] *
                    image[n, c, y + dy_k, x + dx_k] over the K x K window in row-major (dy, dx) order, taps outside
                    the output extent min(image, weights) skipped.
  inv_iter_base     Kernel.cu:137-186 (sigmoid gates from channels K*K, K*K+1, K*K+2 of the fused weights, read
                    with the reference's batch stride of K*K planes).
  inv_iter_kernel   PyTorchDenoiser.cpp:174-179: inv_iter_base, softmax over the first K*K channels, per_pixel_kernel.
"""
import numpy as np

F = np.float32


def per_pixel_kernel(image, weights, k):
    image = np.asarray(image, F)
    weights = np.asarray(weights, F)
    n, c, ih, iw = image.shape
    wh, ww = weights.shape[2:]
    h, w = min(ih, wh), min(iw, ww)
    out = np.zeros((n, c, h, w), F)
    yy, xx = np.mgrid[0:h, 0:w]
    idx = 0
    for dy in range(-(k // 2), k // 2 + 1):
        for dx in range(-(k // 2), k // 2 + 1):
            ky, kx = yy + dy, xx + dx
            ok = (kx >= 0) & (kx < w) & (ky >= 0) & (ky < h)
            tap = image[:, :, np.clip(ky, 0, ih - 1), np.clip(kx, 0, iw - 1)]
            wt = weights[:, idx, :h, :w]
            out += np.where(ok[None, None], wt[:, None] * tap, F(0.0))
            idx += 1
    return out


def _sig(v):
    return (F(1.0) / (F(1.0) + np.exp(-v))).astype(F)


def inv_iter_base(x, lo_pred, lo_x, prev, fused, k, use_previous):
    x, lo_pred, lo_x, fused = (np.asarray(a, F) for a in (x, lo_pred, lo_x, fused))
    n, c, ih, iw = x.shape
    wc, wh, ww = fused.shape[1:]
    lh, lw = lo_x.shape[2:]
    h, w = min(ih, wh), min(iw, ww)
    kd = k * k
    flat = fused.reshape(-1)
    plane = wh * ww
    out = np.zeros((n, c, h, w), F)
    for b in range(n):
        base = b * plane * kd  # the reference's batch stride
        for yy in range(h):
            for xx in range(w):
                lx, ly = xx // 2, yy // 2
                if lx >= lw or ly >= lh:
                    continue
                off = yy * ww + xx
                lfw = _sig(flat[base + kd * plane + off])
                hfw = _sig(flat[base + (kd + 1) * plane + off])
                prw = _sig(flat[base + (kd + 2) * plane + off])
                for ch in range(c):
                    xh, xl, pl = x[b, ch, yy, xx], lo_x[b, ch, ly, lx], lo_pred[b, ch, ly, lx]
                    lf = lfw * pl + (F(1.0) - lfw) * xl
                    if use_previous:
                        # Kernel.cu:176 blends with the double literal 1.0: the product and the sum are in double
                        pv = F(np.asarray(prev, F)[b, ch, yy, xx] * prw)
                        lf = F(np.float64(lf) * (1.0 - np.float64(prw)) + np.float64(pv))
                    out[b, ch, yy, xx] = hfw * (xh - xl + lf) + (F(1.0) - hfw) * lf
    return out


def softmax_channels(fused, kd):
    v = np.asarray(fused, F)[:, :kd].astype(np.float64)
    e = np.exp(v - v.max(1, keepdims=True))
    return (e / e.sum(1, keepdims=True)).astype(F)


def inv_iter_kernel(x, lo_pred, lo_x, prev, fused, k, use_previous):
    base = inv_iter_base(x, lo_pred, lo_x, prev, fused, k, use_previous)
    return per_pixel_kernel(base, softmax_channels(fused, k * k), k)
