"""GPU tests of the kernel-prediction denoiser network (bf16 MFMA convolutions, backward passes, Adam) against a
plain PyTorch float32 reference of the same network.

The reference rounds the weights, the packed inputs and every hidden activation to bfloat16 where the GPU stores
them, so the forward pass differs only by FP32 summation order (and the rare bf16 rounding flips that follow from
it); gradients additionally differ by the bf16 storage of the backward activations.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LOG_CHANNELS = [0, 1, 2, 11, 12, 13]


def _torch():
    import torch

    return torch


def _kpn():
    from cloudrendering_amd import kpn

    return kpn


def _bf16(t):
    return t.to(_torch().bfloat16).float()


def _random_params(seed, scale=1.0):
    kpn = _kpn()
    rng = np.random.default_rng(seed)
    p = np.zeros(kpn.NUM_PARAMS, np.float32)
    w, b = kpn.KpnDenoiser.split(p)
    for l in range(6):
        cin = 15 if l == 0 else 32
        cout = 25 if l == 5 else 32
        bound = scale * np.sqrt(6.0 / (9 * cin))
        w[l, :cout, :cin] = rng.uniform(-bound, bound, (cout, cin, 3, 3))
        b[l, :cout] = rng.uniform(-0.1, 0.1, cout)
    return p


def _inputs(seed, n, h, w):
    torch = _torch()
    g = torch.Generator(device="cpu").manual_seed(seed)
    feats = torch.rand((n, 15, h, w), generator=g)
    feats[:, 6:11] = feats[:, 6:11] * 2.0 - 1.0
    noisy = torch.rand((n, 3, h, w), generator=g) * 2.0
    target = torch.rand((n, 3, h, w), generator=g)
    return feats.cuda(), noisy.cuda(), target.cuda()


def _sanitise(feats):
    """The packing's input guard: the depth-gradient channels 9-10 clamped to +-16 (NaN -> 0), any other
    non-finite feature -> 0."""
    torch = _torch()
    x = feats.clone()
    grad = x[:, 9:11]
    x[:, 9:11] = torch.clamp(torch.nan_to_num(grad, nan=0.0, posinf=16.0, neginf=-16.0), -16.0, 16.0)
    rest = [c for c in range(15) if c not in (9, 10)]
    x[:, rest] = torch.nan_to_num(x[:, rest], nan=0.0, posinf=0.0, neginf=0.0)
    return x


def _reference(params, feats, noisy, target=None):
    """Float32 torch forward (with the GPU's bf16 rounding points), the 5x5 softmax filter and the relative MSE."""
    torch = _torch()
    F = torch.nn.functional
    kpn = _kpn()
    w_np, b_np = kpn.KpnDenoiser.split(params)
    x = _sanitise(feats)
    x[:, LOG_CHANNELS] = torch.log1p(torch.clamp(x[:, LOG_CHANNELS], min=0.0))
    x = _bf16(x)
    x = torch.cat([x, torch.zeros((x.shape[0], 17) + tuple(x.shape[2:]), device=x.device)], 1)
    ws = [_bf16(torch.tensor(w_np[l], device="cuda")).requires_grad_(True) for l in range(6)]
    bs = [torch.tensor(b_np[l], device="cuda").requires_grad_(True) for l in range(6)]
    for l in range(5):
        x = F.relu(F.conv2d(x, ws[l], bs[l], padding=1))
        x = x + (_bf16(x) - x).detach()  # stored as bf16, gradient straight through
    logits = F.conv2d(x, ws[5], bs[5], padding=1)
    k = torch.softmax(logits[:, :25], 1)
    n, _, h, w = noisy.shape
    taps = F.unfold(noisy, 5, padding=2).view(n, 3, 25, h, w)
    pred = (taps * k.unsqueeze(1)).sum(2)
    loss = None
    if target is not None:
        loss = (((pred - target) ** 2) / (target * target + 0.01)).mean()
    return logits, pred, loss, ws, bs


def test_forward_matches_torch(vpt):
    print("case: kpn forward 2x37x45", flush=True)
    torch = _torch()
    kpn = _kpn()
    params = _random_params(1)
    net = kpn.KpnDenoiser(parameters=params)
    feats, noisy, _ = _inputs(2, 2, 37, 45)
    out = net.denoise(feats, noisy)
    logits = net.logits(2, 37, 45).permute(0, 3, 1, 2)
    torch.cuda.synchronize()
    ref_logits, _, _, _, _ = _reference(params, feats, noisy)
    a, b = logits.detach().cpu().numpy(), ref_logits.detach().cpu().numpy()
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel < 1e-2, f"logits relative error {rel}"
    assert np.abs(a[:, 25:]).max() == 0.0
    # the head on the GPU's own logits: softmax, 5x5 filter (zero outside the image), exact up to rounding
    k = torch.softmax(logits[:, :25], 1)
    taps = torch.nn.functional.unfold(noisy, 5, padding=2).view(2, 3, 25, 37, 45)
    pred = (taps * k.unsqueeze(1)).sum(2)
    np.testing.assert_allclose(out.cpu().numpy(), pred.cpu().numpy(), rtol=1e-5, atol=1e-6)


def test_non_finite_features_are_sanitised(vpt):
    """The depth-gradient map is +-inf where an averaged normal lies on a view axis (and NaN where rounding puts the
    normal component above 1); such pixels must not turn the network's logits into NaN."""
    print("case: kpn non-finite features 1x24x40", flush=True)
    torch = _torch()
    kpn = _kpn()
    params = _random_params(5)
    net = kpn.KpnDenoiser(parameters=params)
    feats, noisy, _ = _inputs(6, 1, 24, 40)
    bad = feats.clone()
    bad[0, 9, 3, 4] = float("inf")
    bad[0, 10, 3, 4] = float("-inf")
    bad[0, 9, 10, 20] = float("nan")
    bad[0, 10, 17, 33] = 1e9
    bad[0, 4, 12, 12] = float("nan")
    bad[0, 0, 20, 2] = float("inf")
    out_bad = net.denoise(bad, noisy).clone()
    out_clean = net.denoise(_sanitise(bad), noisy).clone()
    torch.cuda.synchronize()
    assert torch.isfinite(out_bad).all(), "non-finite features leaked into the denoised image"
    assert torch.equal(out_bad, out_clean)
    ref_logits, _, _, _, _ = _reference(params, bad, noisy)
    assert torch.isfinite(ref_logits).all()


def test_one_hot_kernel_is_a_shift(vpt):
    print("case: kpn head one-hot logits", flush=True)
    torch = _torch()
    kpn = _kpn()
    params = np.zeros(kpn.NUM_PARAMS, np.float32)
    _, b = kpn.KpnDenoiser.split(params)
    b[5, 7] = 50.0  # tap 7: (dy, dx) = (-1, 0)
    net = kpn.KpnDenoiser(parameters=params)
    feats, noisy, _ = _inputs(3, 1, 20, 24)
    out = net.denoise(feats, noisy).cpu().numpy()
    ref = np.zeros_like(out)
    ref[:, :, 1:, :] = noisy.cpu().numpy()[:, :, :-1, :]
    np.testing.assert_allclose(out, ref, rtol=1e-5, atol=1e-6)


def test_gradients_match_torch(vpt):
    print("case: kpn gradients 2x33x40", flush=True)
    kpn = _kpn()
    params = _random_params(4)
    net = kpn.KpnDenoiser(parameters=params)
    feats, noisy, target = _inputs(5, 2, 33, 40)
    loss = net.compute_gradients(feats, noisy, target)
    g = net.gradients()
    _, _, ref_loss, ws, bs = _reference(params, feats, noisy, target)
    ref_loss.backward()
    assert abs(loss - ref_loss.item()) <= 2e-3 * abs(ref_loss.item())
    gw, gb = kpn.KpnDenoiser.split(g)
    for l in range(6):
        a, r = gw[l], ws[l].grad.cpu().numpy()
        rel = np.linalg.norm(a - r) / np.linalg.norm(r)
        assert rel < 3e-2, f"layer {l} weight gradient relative error {rel}"
        a, r = gb[l], bs[l].grad.cpu().numpy()
        rel = np.linalg.norm(a - r) / np.linalg.norm(r)
        assert rel < 3e-2, f"layer {l} bias gradient relative error {rel}"
    assert np.abs(gw[0][:, 15:]).max() == 0.0 and np.abs(gw[5][25:]).max() == 0.0 and np.abs(gb[5][25:]).max() == 0.0


def test_adam_step_matches_formula(vpt):
    print("case: kpn adam", flush=True)
    torch = _torch()
    kpn = _kpn()
    params = _random_params(6)
    net = kpn.KpnDenoiser(parameters=params)
    feats, noisy, target = _inputs(7, 1, 24, 24)
    p = net.parameters().astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    lr = 1e-3
    for t in range(1, 4):
        net.compute_gradients(feats, noisy, target)
        g = net.gradients().astype(np.float64)
        net.adam_step(lr)
        torch.cuda.synchronize()
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        p = p - lr * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
        got = net.parameters()
        np.testing.assert_allclose(got, p, rtol=1e-5, atol=1e-7)
        p = got.astype(np.float64)


def test_training_reduces_loss(vpt):
    print("case: kpn training 60 steps", flush=True)
    torch = _torch()
    kpn = _kpn()
    net = kpn.KpnDenoiser(seed=3)
    g = torch.Generator(device="cpu").manual_seed(8)
    h = w = 48
    yy, xx = torch.meshgrid(torch.linspace(0, 1, h), torch.linspace(0, 1, w), indexing="ij")
    clean = torch.stack([0.5 + 0.4 * torch.sin(6 * xx + c) * torch.cos(4 * yy) for c in range(3)]).unsqueeze(0)
    clean = clean.repeat(4, 1, 1, 1)
    noisy = clean + 0.15 * torch.randn(clean.shape, generator=g)
    feats = torch.zeros((4, 15, h, w))
    feats[:, 0:3] = noisy
    feats[:, 3:6] = 0.9
    feats[:, 11:14] = 0.15 ** 2
    feats, noisy, clean = feats.cuda(), noisy.cuda(), clean.cuda()
    first = net.compute_gradients(feats, noisy, clean)
    losses = [net.train_step(feats, noisy, clean, 2e-3) for _ in range(60)]
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.6 * first, f"loss {first} -> {losses[-1]}"
