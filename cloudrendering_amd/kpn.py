"""Kernel-prediction denoiser (network, training and inference on the GPU through libcrvpt's crvpt_kpn_* entry
points), and the renderer-side helpers that produce its inputs.

The network (six bf16 3x3 convolutions on the matrix cores, 15 -> 32 x 5 -> 25 channels) predicts per pixel a 5x5
filter (softmax of 25 logits) that is applied to the noisy radiance, as the reference's per_pixel_kernel operator
applies the kernels of its TorchScript denoiser (modules/kpn_module/src/Kernel.cu:38-83, PyTorchDenoiser.cpp:410-443).
Tensors are float32 on the GPU; torch only hands the buffers over.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import CrvptError, call

NUM_PARAMS = 55488
NUM_INPUTS = 15
# Input feature channels, in order.
FEATURE_NAMES = [
    "radiance r", "radiance g", "radiance b", "albedo r", "albedo g", "albedo b", "depth", "depth std",
    "transmittance", "gradient x", "gradient y", "variance r", "variance g", "variance b", "density",
]


def _torch():
    import torch  # noqa: PLC0415

    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(t):
    return ctypes.c_void_p(_torch().cuda.current_stream(t.device).cuda_stream)


def _check(t, shape_tail, name):
    torch = _torch()
    if t.device.type != "cuda" or t.dtype != torch.float32 or t.dim() != 4 or tuple(t.shape[1:2]) != shape_tail:
        raise CrvptError(f"kpn: {name} must be a float32 [N, {shape_tail[0]}, H, W] tensor on the GPU")
    return t.contiguous()


class KpnDenoiser:
    """The library's single KPN model (one per process)."""

    def __init__(self, seed: int = 0, parameters=None):
        if parameters is None:
            call("crvpt_kpn_init", int(seed))
        else:
            self.set_parameters(parameters)

    # parameters: weights [6, 32, 32, 3, 3] then biases [6, 32], flattened
    def parameters(self) -> np.ndarray:
        out = np.zeros(NUM_PARAMS, np.float32)
        call("crvpt_kpn_get_parameters", out.ctypes.data_as(ctypes.c_void_p), NUM_PARAMS)
        return out

    def set_parameters(self, params) -> None:
        p = np.ascontiguousarray(np.asarray(params, np.float32).reshape(-1))
        call("crvpt_kpn_set_parameters", p.ctypes.data_as(ctypes.c_void_p), int(p.size))

    def gradients(self) -> np.ndarray:
        out = np.zeros(NUM_PARAMS, np.float32)
        call("crvpt_kpn_get_gradients", out.ctypes.data_as(ctypes.c_void_p), NUM_PARAMS)
        return out

    @staticmethod
    def split(params):
        """(weights [6, 32, 32, 3, 3], biases [6, 32]) views of a flat parameter vector."""
        p = np.asarray(params).reshape(-1)
        return p[:6 * 9216].reshape(6, 32, 32, 3, 3), p[6 * 9216:].reshape(6, 32)

    def compute_gradients(self, features, noisy, target) -> float:
        f = _check(features, (NUM_INPUTS,), "features")
        x = _check(noisy, (3,), "noisy")
        t = _check(target, (3,), "target")
        n, _, h, w = (int(v) for v in x.shape)
        loss = ctypes.c_float()
        call("crvpt_kpn_compute_gradients", _ptr(f), _ptr(x), _ptr(t), n, h, w, ctypes.byref(loss), _stream(x))
        return float(loss.value)

    def adam_step(self, lr: float) -> None:
        call("crvpt_kpn_adam_step", float(lr), ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream))

    def train_step(self, features, noisy, target, lr: float) -> float:
        f = _check(features, (NUM_INPUTS,), "features")
        x = _check(noisy, (3,), "noisy")
        t = _check(target, (3,), "target")
        n, _, h, w = (int(v) for v in x.shape)
        loss = ctypes.c_float()
        call("crvpt_kpn_train_step", _ptr(f), _ptr(x), _ptr(t), n, h, w, float(lr), ctypes.byref(loss), _stream(x))
        return float(loss.value)

    def denoise(self, features, noisy):
        torch = _torch()
        f = _check(features, (NUM_INPUTS,), "features")
        x = _check(noisy, (3,), "noisy")
        out = torch.empty_like(x)
        n, _, h, w = (int(v) for v in x.shape)
        call("crvpt_kpn_denoise", _ptr(f), _ptr(x), _ptr(out), n, h, w, _stream(x))
        return out

    def logits(self, n: int, h: int, w: int, device="cuda"):
        torch = _torch()
        out = torch.empty((n, h, w, 32), dtype=torch.float32, device=device)
        call("crvpt_kpn_get_logits", _ptr(out), int(n), int(h), int(w), _stream(out))
        return out


def render_inputs(vpt, height: int, width: int, spp: int = 4):
    """Renders spp samples per pixel of the renderer's current scene and returns (features [15, H, W], noisy radiance
    [3, H, W]) on the GPU: the radiance mean and its per-pixel sample variance (the Radiance Variance map, scaled to
    the unbiased estimate), and the reference's first-event feature maps (Albedo, Depth, First X coverage,
    Depth (nabla), Density)."""
    torch = _torch()
    vpt.set_use_feature_maps(["Albedo", "Depth", "First X", "Depth (nabla)", "Density", "Radiance Variance"])
    probe = torch.zeros((4, height, width), dtype=torch.float32, device="cuda")
    mu = vpt.render_frame(probe, spp)[:3]
    var = vpt.get_feature_map(probe, 21)[:3] * (spp / max(spp - 1, 1))
    alb = vpt.get_feature_map(probe, 17)
    dep = vpt.get_feature_map(probe, 6)
    fx = vpt.get_feature_map(probe, 1)
    nab = vpt.get_feature_map(probe, 10)
    den = vpt.get_feature_map(probe, 12)
    feats = torch.cat([mu, alb[:3], dep[0:1], dep[1:2], 1.0 - fx[3:4], nab[0:2], var, den[0:1]], 0)
    return feats.contiguous(), mu.contiguous()


def rel_mse(x, ref) -> float:
    """Mean over pixels and channels of (x - r)^2 / (r^2 + 0.01)."""
    return float((((x - ref) ** 2) / (ref * ref + 0.01)).mean())
