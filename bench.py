"""Throughput benchmark of the gfx950 volumetric path tracer on its headline workload.

Default workload "c2_sparse512" (one GPU): a NanoVDB FP32 sparse grid of a 512^3 domain holding a seeded fBm cloud
(5 octaves, seed 0, lacunarity 2, gain 0.5, base frequency 4) thresholded by smoothstep(0.645, 0.795) so that about
12 % of the voxels are active in 8^3 leaves, 16^3-voxel majorant supervoxels, extinction 1024, albedo 0.9, the
reference's default sky, 1024 x 1024 pinhole camera at the reference module's default pose, 256 samples per pixel,
at most 128 scattering events per path. One step renders the image once with every one of the 12 combinations
(delta, ratio, residual ratio, decomposition tracking) x g in {0, 0.5, 0.85}: 12 x 268 M paths.

Workload "c1_dense128": dense 128^3 FP32 fBm (smoothstep(0.45, 0.6)) filling the unit cube, next-event tracking with
a constant white environment image and one directional sun, g = 0.85, 256 x 256, 64 spp per step, at most 64
scattering events per path.

Workload "c3_batch": the batched render queue, 64 requests (dense 64^3 / 128^3 / 256^3 and NanoVDB 512^3 / 1024^3
volumes, image side and spp log-uniform in 128-1024 and 16-1024, random camera orbits and sun directions) traced by
one persistent launch per step with the autotuned launch configuration; prints its own JSON line.

Workload "c4_kpn": the kernel-prediction denoiser at the C4 scale: 4096 training crops rendered by the tracer (4-spp
and 4-256-spp inputs with feature maps, 1024-spp targets in one batched launch), training on the GPU, the
equal-time comparison with EAW, SVGF and plain tracking at the 4-spp budget, and the wall time to relMSE 1e-3
against plain tracking, all judged by a 16k-spp render of an unseen cloud; prints its own JSON line.

Workload "c5_fp16_2048": the large sparse production frame: NanoVDB with Fp16 leaves over a 2048^3 domain (6-octave
anvil fBm, about 5 % active), 1920 x 1080, 1024 spp, residual ratio tracking, at most 256 scattering events, plus in
the same loop per frame an 8-spp render with feature maps (training-pair input), kernel-prediction training steps on
crops of the pairs and the network's inference on the full frame; prints its own JSON line.

Prints one JSON line: Mpaths/s over the timed steps, the roofline of the dominant kernel (algorithmic bytes from
device work counters of an instrumented pass with identical seeds, divided by the HIP-event time of that kernel in
the timed passes), and the CPU oracle timed on a bounded sample of the same scene on the host cores. The default
workload's line also carries the metric's second half, measured after the timed steps: the c4_kpn workload's wall
time to relMSE 1e-3 (KPN-denoised against plain tracking), under "kpn_wall_time_to_rel_mse_1e-3".
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0

C2 = {"max_bounces": 128, "grid": 512, "edge0": 0.645, "edge1": 0.795, "width": 1024, "height": 1024, "spp": 256,
      "modes": [(0, "delta tracking"), (2, "ratio tracking"), (4, "residual ratio tracking"),
                (3, "decomposition tracking")], "gs": [0.0, 0.5, 0.85]}
C1 = {"max_bounces": 64, "grid": 128, "width": 256, "height": 256, "spp": 64}
MODE_KERNELS = {0: "k_pathPool<DELTA_TRACKING, NanoAccessT<false, false>, false>", 2: "k_pathPool<RATIO_TRACKING, NanoAccessT<false, false>, false>",
                4: "k_pathPool<RESIDUAL_RATIO_TRACKING, NanoAccessT<false, false>, false>",
                3: "k_pathPool<DECOMPOSITION_TRACKING, NanoAccessT<false, false>, false>"}


def setup_c1(vpt, torch):
    vpt.set_max_grid_extent(0.5)
    vpt.generate_fbm_cloud(C1["grid"], C1["grid"], C1["grid"], seed=0, octaves=5, lacunarity=2.0, gain=0.5,
                           base_frequency=4.0, edge0=0.45, edge1=0.6)
    vpt.set_vpt_mode_from_name("Next Event Tracking")
    vpt.set_grid_interpolation_type(1)
    vpt.set_extinction_base([1.0, 1.0, 1.0])
    vpt.set_extinction_scale(1024.0)
    vpt.set_scattering_albedo([0.9, 0.9, 0.9])
    vpt.set_phase_g(0.85)
    vpt.load_environment_map_from_tensor(torch.ones((3, 16, 32), dtype=torch.float32))
    vpt.set_use_lights(True)
    vpt.clear_lights()
    vpt.add_light()
    vpt.set_light_property(0, "type", "Directional Light")
    vpt.set_light_property(0, "space", "World")
    # direction norm(0.5826, 0.7660, 0.2717); length 100 so the distance-limited shadow ray reaches the box exit
    d = [0.5826, 0.7660, 0.2717]
    n = math.sqrt(sum(v * v for v in d))
    vpt.set_light_property(0, "position", " ".join(f"{100.0 * v / n:.6f}" for v in d))
    vpt.set_light_property(0, "color", "1 0.961538462 0.884615385")
    vpt.set_light_property(0, "intensity", "2.6")
    vpt.set_max_bounces(C1["max_bounces"])


def setup_c2(vpt):
    g = C2["grid"]
    vpt.generate_fbm_cloud(g, g, g, seed=0, octaves=5, lacunarity=2.0, gain=0.5, base_frequency=4.0,
                           edge0=C2["edge0"], edge1=C2["edge1"])
    vpt.set_grid_interpolation_type(1)
    vpt.set_extinction_base([1.0, 1.0, 1.0])
    vpt.set_extinction_scale(1024.0)
    vpt.set_scattering_albedo([0.9, 0.9, 0.9])
    vpt.set_use_builtin_environment_map("Default")
    vpt.set_use_lights(False)
    vpt.set_use_sparse_grid(True)
    vpt.set_launch_config(16, 0, 0)
    vpt.set_super_voxel_size(16)
    vpt.set_max_bounces(C2["max_bounces"])


C3 = {"requests": 64, "max_bounces": 128,
      "volumes": [("dense", 64), ("dense", 128), ("dense", 256), ("sparse", 512), ("sparse", 1024)],
      "res": (128, 1024), "spp": (16, 1024)}


def c3_requests():
    """The 64 requests of the batched-queue workload, drawn with request index i as the seed: volume kind uniform
    over dense {64^3, 128^3, 256^3} and sparse {512^3, 1024^3}, image side and spp log-uniform, camera on an orbit
    around the volume, sun (a directional light) from a random direction above the horizon."""
    import numpy as np

    reqs = []
    for i in range(C3["requests"]):
        rng = np.random.default_rng(i)
        kind = int(rng.integers(len(C3["volumes"])))
        side = int(round(math.exp(rng.uniform(math.log(C3["res"][0]), math.log(C3["res"][1])))))
        spp = int(round(math.exp(rng.uniform(math.log(C3["spp"][0]), math.log(C3["spp"][1])))))
        az = rng.uniform(0.0, 2.0 * math.pi)
        el = rng.uniform(-math.pi / 6.0, math.pi / 3.0)
        cam = [0.8 * math.cos(el) * math.sin(az), 0.8 * math.sin(el), 0.8 * math.cos(el) * math.cos(az)]
        sun = rng.normal(size=3)
        sun[1] = abs(sun[1]) + 0.2
        sun = sun / np.linalg.norm(sun)
        reqs.append({"index": i, "kind": kind, "side": side, "spp": spp, "camera": cam, "sun": sun.tolist()})
    return reqs


def c3_settings(vpt):
    """Scene settings shared by every request of the batched-queue workload: next-event tracking with the default
    sky and one directional sun, extinction 1024, albedo 0.9, g 0.5, at most 128 scattering events per path."""
    vpt.set_vpt_mode(5)
    vpt.set_grid_interpolation_type(1)
    vpt.set_extinction_base([1.0, 1.0, 1.0])
    vpt.set_extinction_scale(1024.0)
    vpt.set_scattering_albedo([0.9, 0.9, 0.9])
    vpt.set_phase_g(0.5)
    vpt.set_use_builtin_environment_map("Default")
    vpt.set_max_bounces(C3["max_bounces"])
    vpt.set_use_lights(True)
    vpt.clear_lights()
    vpt.add_light()
    vpt.set_light_property(0, "type", "Directional Light")
    vpt.set_light_property(0, "space", "World")
    vpt.set_light_property(0, "color", "1 0.961538462 0.884615385")
    vpt.set_light_property(0, "intensity", "2.6")


def c3_volume(vpt, kind):
    """The volume of request kind `kind`: one seeded fBm volume per kind (dense: smoothstep(0.45, 0.6), ~30 %
    nonzero; sparse: smoothstep(0.645, 0.795), ~12 % active in 8^3 leaves)."""
    rep, n = C3["volumes"][kind]
    sparse = rep == "sparse"
    e0, e1 = (0.645, 0.795) if sparse else (0.45, 0.6)
    vpt.generate_fbm_cloud(n, n, n, seed=100 + kind, octaves=5, lacunarity=2.0, gain=0.5, base_frequency=4.0,
                           edge0=e0, edge1=e1)
    vpt.set_use_sparse_grid(sparse)
    vpt.set_super_voxel_size(16 if sparse else 8)


def c3_view(vpt, q):
    """Camera, sun direction and seed offset of request q."""
    vpt.set_camera_position(q["camera"])
    vpt.set_camera_target([0.0, 0.0, 0.0])
    vpt.set_light_property(0, "position", " ".join(f"{100.0 * v:.6f}" for v in q["sun"]))
    vpt.set_seed_offset(q["index"])


def c3_spp(q, spp_divisor=1):
    return max(1, q["spp"] // spp_divisor)


def setup_c3(vpt, torch, spp_divisor=1):
    """Queues the batched-queue workload (c3_requests), volume by volume; spp_divisor > 1 queues every request with
    its spp divided (at least 1)."""
    reqs = c3_requests()
    vpt.batch_clear()
    c3_settings(vpt)
    for kind in range(len(C3["volumes"])):
        mine = [q for q in reqs if q["kind"] == kind]
        if not mine:
            continue
        c3_volume(vpt, kind)
        for q in mine:
            c3_view(vpt, q)
            vpt.batch_add_request(torch.zeros((3, q["side"], q["side"]), dtype=torch.float32, device="cuda"),
                                  c3_spp(q, spp_divisor))
    return reqs


def run_c3(args, vpt, torch):
    reqs = setup_c3(vpt, torch)
    t = time.perf_counter()
    cfg = vpt.batch_autotune(16)
    tune_s = time.perf_counter() - t
    print(f"autotune {cfg} in {tune_s:.1f} s", file=sys.stderr, flush=True)
    for k in range(args.warmup):
        vpt.batch_render(k * 1024)
    torch.cuda.synchronize()
    ms = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        vpt.batch_render((args.warmup + k) * 1024)
        st = vpt.batch_get_stats()
        ms.append(st["kernel_ms"])
        print(f"step {k}: batch {st['kernel_ms']:.1f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    paths = st["num_paths"]
    value = paths / (wall / args.steps) / 1e6
    line = {
        "metric": "Mpaths/s on 1 MI355X", "value": round(value, 3), "unit": "Mpaths/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded fBm procedural clouds generated on the GPU)",
        "config": {"workload": "c3_batch", "requests": len(reqs), "paths_per_step": paths,
                   "volumes": "dense 64^3/128^3/256^3, NanoVDB fp32 512^3/1024^3 (one seeded fBm volume per kind)",
                   "resolution": "log-uniform side 128-1024", "spp": "log-uniform 16-1024",
                   "estimator": "next event tracking", "launch": "one persistent launch per step (HIP graph)",
                   "autotuned": cfg, "autotune_s": round(tune_s, 2), "max_bounces": C3["max_bounces"]},
        "batch_kernel_ms": [round(v, 3) for v in ms],
    }
    print(json.dumps(line), flush=True)


C4 = {"crops": 4096, "crop": 128, "clouds": 16, "input_spp": 4, "mixed_input_spp": [4, 16, 32, 48, 64, 256],
      "eval_input_spp": [4, 16, 32, 48, 64, 128, 256], "target_spp": 1024, "batch": 16, "train_steps": 6000, "lr": 1e-3,
      "eval_res": 1024, "ref_spp": 16384, "plain_spp": [4, 16, 64, 256, 1024, 4096], "max_bounces": 128,
      "eaw_iterations": 4, "svgf_max_frames": 256}


def _c4_scene(vpt, seed):
    """Dense 128^3 seeded fBm cloud (smoothstep(0.45, 0.6)), delta tracking, extinction 1024, albedo 0.9, g 0.5,
    default sky, at most 128 scattering events."""
    vpt.generate_fbm_cloud(128, 128, 128, seed=seed, octaves=5, lacunarity=2.0, gain=0.5, base_frequency=4.0,
                           edge0=0.45, edge1=0.6)
    vpt.set_use_sparse_grid(False)
    vpt.set_vpt_mode(0)
    vpt.set_grid_interpolation_type(1)
    vpt.set_extinction_base([1.0, 1.0, 1.0])
    vpt.set_extinction_scale(1024.0)
    vpt.set_scattering_albedo([0.9, 0.9, 0.9])
    vpt.set_phase_g(0.5)
    vpt.set_use_builtin_environment_map("Default")
    vpt.set_use_lights(False)
    vpt.set_max_bounces(C4["max_bounces"])


def _c4_camera(vpt, rng):
    az = rng.uniform(0.0, 2.0 * math.pi)
    el = rng.uniform(-math.pi / 8.0, math.pi / 3.0)
    d = rng.uniform(0.45, 0.8)
    vpt.set_camera_position([d * math.cos(el) * math.sin(az), d * math.sin(el), d * math.cos(el) * math.cos(az)])
    vpt.set_camera_target([0.0, 0.0, 0.0])


def _train(torch, net, feats, noisy, target, steps, gen, log):
    """`steps` Adam steps on random batches of the pairs (learning rate / 10 for the last fifth)."""
    losses = []
    n = feats.shape[0]
    for s_ in range(steps):
        idx = torch.randint(0, n, (C4["batch"],), generator=gen).cuda()
        lr = C4["lr"] * (0.1 if s_ > 0.8 * steps else 1.0)
        losses.append(net.train_step(feats[idx], noisy[idx], target[idx], lr))
        if s_ % 1000 == 0:
            print(f"{log} step {s_}: loss {losses[-1]:.5f}", file=sys.stderr, flush=True)
    return losses


def run_c4(args, vpt, torch, emit=True):
    """Kernel-prediction denoiser at the C4 scale: 4096 seeded 128^2 crops of 16 fBm clouds, each with a 1024-spp
    target (all targets in one batched launch) and two inputs with the tracer's feature maps: 4 spp (network A, the
    4-spp -> 1024-spp training of the configuration) and 4, 16, 32, 48, 64 or 256 spp in turn (network B, the same
    crops and targets). Evaluation on an unseen cloud and camera at 1024^2, judged against a 16k-spp render:
      * equal time: network A on a 4-spp render against EAW (4 a-trous iterations on the same 4-spp render), SVGF (1-spp
        frames of the static view accumulated by its temporal filter for the same wall time) and plain tracking at
        the same wall time (interpolated on the plain curve);
      * the headline: wall time to relMSE 1e-3, network B (render at 4-256 input spp + features + inference) against
        plain tracking at increasing spp."""
    import numpy as np

    from cloudrendering_amd import kpn

    crops = args.kpn_crops or C4["crops"]
    cs = C4["crop"]
    # the renderer's default launch configuration (majorant supervoxels, workgroups, refill interval), whatever a
    # workload run before this one in the same process set
    vpt.set_launch_config(0, 0, 0)
    featsA = torch.zeros((crops, 15, cs, cs), dtype=torch.float32, device="cuda")
    noisyA = torch.zeros((crops, 3, cs, cs), dtype=torch.float32, device="cuda")
    featsB = torch.zeros((crops, 15, cs, cs), dtype=torch.float32, device="cuda")
    noisyB = torch.zeros((crops, 3, cs, cs), dtype=torch.float32, device="cuda")
    t0 = time.perf_counter()
    vpt.batch_clear()
    per_cloud = (crops + C4["clouds"] - 1) // C4["clouds"]
    targets = []
    for c in range(crops):
        if c % per_cloud == 0:
            _c4_scene(vpt, 1 + c // per_cloud)
            print(f"pairs: cloud {1 + c // per_cloud}, crop {c}", file=sys.stderr, flush=True)
        rng = np.random.default_rng(10_000 + c)
        _c4_camera(vpt, rng)
        vpt.set_use_feature_maps([])
        vpt.set_seed_offset(3 * c + 2)
        targets.append(vpt.batch_add_request(torch.zeros((3, cs, cs), device="cuda"), C4["target_spp"])[1])
        vpt.set_seed_offset(3 * c)
        f, x = kpn.render_inputs(vpt, cs, cs, C4["input_spp"])
        featsA[c] = f
        noisyA[c] = x
        vpt.set_seed_offset(3 * c + 1)
        f, x = kpn.render_inputs(vpt, cs, cs, C4["mixed_input_spp"][c % len(C4["mixed_input_spp"])])
        featsB[c] = f
        noisyB[c] = x
    vpt.set_use_feature_maps([])
    vpt.batch_set_launch_config(0, 256, 16, 64)
    vpt.batch_render()
    torch.cuda.synchronize()
    target = torch.stack(targets)
    del targets
    vpt.batch_clear()
    gen_s = time.perf_counter() - t0
    print(f"pairs: {crops} crops in {gen_s:.1f} s", file=sys.stderr, flush=True)

    steps = args.kpn_steps or C4["train_steps"]
    g = torch.Generator(device="cpu").manual_seed(0)
    netA = kpn.KpnDenoiser(seed=0)
    netB = kpn.KpnDenoiser(seed=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lossA = _train(torch, netA, featsA, noisyA, target, steps, g, "train A (4 spp)")
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    lossB = _train(torch, netB, featsB, noisyB, target, steps, g, "train B (4-256 spp)")
    torch.cuda.synchronize()
    del featsA, noisyA, featsB, noisyB, target

    # evaluation on an unseen cloud and camera
    res = args.kpn_eval_res or C4["eval_res"]
    ref_spp = args.kpn_ref_spp or C4["ref_spp"]
    _c4_scene(vpt, 999)
    _c4_camera(vpt, np.random.default_rng(999))
    vpt.set_seed_offset(7777)
    vpt.batch_clear()
    _, ref = vpt.batch_add_request(torch.zeros((3, res, res), device="cuda"), ref_spp)
    vpt.batch_set_launch_config(0, 256, 16, 256)
    vpt.batch_render()
    torch.cuda.synchronize()
    vpt.batch_clear()
    probe = torch.zeros((3, res, res), dtype=torch.float32, device="cuda")

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t

    plain = []
    for spp in C4["plain_spp"]:
        vpt.set_seed_offset(100 + spp)
        img, sec = timed(lambda: vpt.render_frame(probe, spp))
        plain.append({"spp": spp, "seconds": round(sec, 4), "rel_mse": kpn.rel_mse(img, ref)})

    def plain_at(seconds):
        """Plain-tracking relMSE after `seconds`, interpolated log-log on the measured curve."""
        for a, b in zip(plain, plain[1:]):
            if a["seconds"] <= seconds <= b["seconds"]:
                u = (math.log(seconds) - math.log(a["seconds"])) / (math.log(b["seconds"]) - math.log(a["seconds"]))
                return math.exp(math.log(a["rel_mse"]) + u * (math.log(b["rel_mse"]) - math.log(a["rel_mse"])))
        return None

    def time_to(err):
        """Plain-tracking seconds to reach err, interpolated log-log between the measured spp."""
        for a, b in zip(plain, plain[1:]):
            if a["rel_mse"] >= err >= b["rel_mse"]:
                u = (math.log(a["rel_mse"]) - math.log(err)) / (math.log(a["rel_mse"]) - math.log(b["rel_mse"]))
                return math.exp(math.log(a["seconds"]) + u * (math.log(b["seconds"]) - math.log(a["seconds"])))
        return None

    def kpn_run(net, spp, seed):
        vpt.set_seed_offset(seed)

        def go():
            f, x = kpn.render_inputs(vpt, res, res, spp)
            return net.denoise(f.unsqueeze(0), x.unsqueeze(0))[0], x
        (den, x), sec = timed(go)
        vpt.set_use_feature_maps([])
        return den, x, sec

    # equal time at the 4-spp budget: network A, EAW, SVGF, plain
    denA, xA, secA = kpn_run(netA, C4["input_spp"], 55)
    vpt.set_denoiser("EAW")
    vpt.set_denoiser_property("maxNumIterations", str(C4["eaw_iterations"]))
    vpt.set_seed_offset(56)
    eaw, sec_eaw = timed(lambda: vpt.render_frame(probe, C4["input_spp"]))
    vpt.set_denoiser("SVGF")
    vpt.set_denoiser_property("maxNumIterations", "5")
    vpt.set_seed_offset(57)
    frames, sec_svgf = 0, 0.0
    while sec_svgf < secA and frames < C4["svgf_max_frames"]:
        svgf, sec = timed(lambda: vpt.render_frame(probe, 1))
        sec_svgf += sec
        frames += 1
    vpt.set_denoiser("None")
    vpt.set_use_feature_maps([])
    equal_time = {
        "budget_s": round(secA, 4),
        "kpn_4spp": {"seconds": round(secA, 4), "rel_mse": kpn.rel_mse(denA, ref), "noisy_rel_mse": kpn.rel_mse(xA, ref)},
        "eaw_4spp": {"seconds": round(sec_eaw, 4), "iterations": C4["eaw_iterations"], "rel_mse": kpn.rel_mse(eaw, ref)},
        "svgf": {"seconds": round(sec_svgf, 4), "frames_1spp": frames, "rel_mse": kpn.rel_mse(svgf, ref)},
        "plain": {"seconds": round(secA, 4), "rel_mse": plain_at(secA)},
    }
    print(f"equal time: {equal_time}", file=sys.stderr, flush=True)

    # headline: wall time to relMSE 1e-3, network B at increasing input spp against plain tracking
    denoised = []
    for spp in C4["eval_input_spp"]:
        den, x, sec = kpn_run(netB, spp, 500 + spp)
        d = {"input_spp": spp, "seconds": round(sec, 4), "rel_mse": kpn.rel_mse(den, ref),
             "noisy_rel_mse": kpn.rel_mse(x, ref)}
        d["plain_seconds_at_same_rel_mse"] = time_to(d["rel_mse"])
        d["speedup_at_same_rel_mse"] = (d["plain_seconds_at_same_rel_mse"] / d["seconds"]
                                        if d["plain_seconds_at_same_rel_mse"] else None)
        denoised.append(d)

    # Denoiser inference throughput on the full evaluation frame (HIP events on the stream the network runs on):
    # algorithmic FLOPs per pixel = 2 * 9 * (15 * 32 + 4 * 32 * 32 + 32 * 25) for the six 3x3 convolutions.
    f, x = kpn.render_inputs(vpt, res, res, 4)
    vpt.set_use_feature_maps([])
    f, x = f.unsqueeze(0), x.unsqueeze(0)
    netB.denoise(f, x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record()
    for _ in range(reps):
        netB.denoise(f, x)
    e1.record()
    torch.cuda.synchronize()
    inf_ms = e0.elapsed_time(e1) / reps
    conv_flops = 2 * 9 * (15 * 32 + 4 * 32 * 32 + 32 * 25) * res * res
    inf_tflops = conv_flops / (inf_ms * 1e-3) / 1e12
    inference = {"pixels": res * res, "ms_per_frame": round(inf_ms, 4), "conv_flops_per_frame": conv_flops,
                 "roofline": {"bound": "mfma", "achieved": round(inf_tflops, 2), "peak": 2500.0, "unit": "TFLOP/s",
                              "frac": round(inf_tflops / 2500.0, 5),
                              "what": "whole denoise call (feature packing, six bf16 MFMA convolutions, softmax "
                                      "5x5 filter) against the dense bf16 MFMA peak"}}

    plain_1e3 = time_to(1e-3)
    reached = [d for d in denoised if d["rel_mse"] <= 1e-3]
    kpn_1e3 = min(d["seconds"] for d in reached) if reached else None
    # the same log-log interpolation the plain-tracking side uses, between the measured denoised points bracketing
    # 1e-3 (reported beside the measured value, which is the first evaluated input spp that reaches 1e-3)
    kpn_interp = None
    for a, b in zip(denoised, denoised[1:]):
        if a["rel_mse"] >= 1e-3 >= b["rel_mse"]:
            u = (math.log(a["rel_mse"]) - math.log(1e-3)) / (math.log(a["rel_mse"]) - math.log(b["rel_mse"]))
            kpn_interp = math.exp(math.log(a["seconds"]) + u * (math.log(b["seconds"]) - math.log(a["seconds"])))
            break
    line = {
        "metric": "wall time to relMSE 1e-3: KPN-denoised vs plain tracking", "unit": "s",
        "value": round(kpn_1e3, 4) if kpn_1e3 else None, "higher_is_better": False, "n_gpus": 1,
        "dtype": "bf16 (network, FP32 accumulation) / f32 (tracer)", "data": "synthetic (seeded fBm clouds)",
        "config": {"workload": "c4_kpn", "crops": crops, "crop": cs, "clouds": C4["clouds"],
                   "network_a_input_spp": C4["input_spp"], "network_b_input_spp": C4["mixed_input_spp"],
                   "eval_input_spp": C4["eval_input_spp"], "target_spp": C4["target_spp"], "train_steps": steps,
                   "batch": C4["batch"], "eval_res": res, "ref_spp": ref_spp, "estimator": "delta tracking",
                   "volume": "dense 128^3 fBm per cloud (seeds 1-16 training, 999 evaluation)"},
        "pair_generation_s": round(gen_s, 2), "train_s_per_network": round(train_s, 2),
        "final_train_loss": {"a": lossA[-1], "b": lossB[-1]}, "train_steps_per_s": round(steps / train_s, 1),
        "equal_time_4spp": equal_time, "inference": inference,
        "kpn": denoised, "plain": plain, "plain_seconds_to_1e-3": plain_1e3,
        "kpn_seconds_to_1e-3_interpolated": round(kpn_interp, 4) if kpn_interp else None,
        "speedup_to_1e-3": (plain_1e3 / kpn_1e3) if (plain_1e3 and kpn_1e3) else None,
    }
    if emit:
        print(json.dumps(line), flush=True)
    return line


C5 = {"grid": 2048, "octaves": 6, "edge0": 0.78, "edge1": 0.93, "width": 1920, "height": 1080, "spp": 1024,
      "max_bounces": 256, "pair_spp": 8, "crop": 128, "batch": 16, "train_steps": 100, "lr": 1e-3}


def setup_c5(vpt):
    """Large sparse production frame: NanoVDB with Fp16 leaves over a 2048^3 domain holding a 6-octave seeded fBm
    with the anvil height profile, smoothstep(0.78, 0.93) (about 5 % of the voxels active), 16^3 supervoxels,
    residual ratio tracking with the per-supervoxel control extinction, extinction 1024, albedo 0.9, g 0.5, default
    sky, at most 256 scattering events per path."""
    g = C5["grid"]
    print(f"c5: generating the {g}^3 fBm field", file=sys.stderr, flush=True)
    vpt.set_sparse_value_type(1)
    vpt.generate_fbm_cloud(g, g, g, seed=0, octaves=C5["octaves"], lacunarity=2.0, gain=0.5, base_frequency=4.0,
                           edge0=C5["edge0"], edge1=C5["edge1"], height_falloff=True)
    vpt.set_use_sparse_grid(True)
    vpt.set_super_voxel_size(16)
    vpt.set_vpt_mode(4)
    vpt.set_grid_interpolation_type(1)
    vpt.set_extinction_base([1.0, 1.0, 1.0])
    vpt.set_extinction_scale(1024.0)
    vpt.set_scattering_albedo([0.9, 0.9, 0.9])
    vpt.set_phase_g(0.5)
    vpt.set_use_builtin_environment_map("Default")
    vpt.set_use_lights(False)
    vpt.set_max_bounces(C5["max_bounces"])


def _c5_camera(vpt, k):
    az = 0.35 * k
    el = 0.15
    vpt.set_camera_position([0.8 * math.cos(el) * math.sin(az), 0.8 * math.sin(el), 0.8 * math.cos(el) * math.cos(az)])
    vpt.set_camera_target([0.0, 0.0, 0.0])


def run_c5(args, vpt, torch):
    """One production frame per step: the 1920x1080, 1024-spp residual-ratio render of the 2048^3 Fp16 grid, then
    in the same loop an 8-spp render of the same view with the denoiser's feature maps (the low-spp half of a
    training pair whose high-spp half is the frame itself), training steps of the kernel-prediction network on
    128^2 crops of every pair so far, and the network's inference on the full 8-spp frame. The camera moves on an
    orbit from frame to frame."""
    import ctypes

    import numpy as np

    from cloudrendering_amd import kpn

    t0 = time.perf_counter()
    setup_c5(vpt)
    W, H, spp, pspp = C5["width"], C5["height"], args.spp or C5["spp"], C5["pair_spp"]
    target = torch.zeros((3, H, W), dtype=torch.float32, device="cuda")
    _c5_camera(vpt, 0)
    print(f"c5: field generated in {time.perf_counter() - t0:.1f} s; building the Fp16 grid and the tables",
          file=sys.stderr, flush=True)
    vpt.render_frame(torch.zeros((3, 64, 64), dtype=torch.float32, device="cuda"), 1)  # builds grid and tables
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    size = ctypes.c_uint64()
    vpt.call("crvpt_get_nanovdb_buffer", None, 0, ctypes.byref(size))
    nano = {"bytes": int(size.value), "leaf_bytes": 1120}
    print(f"c5 setup (generate 2048^3, Fp16 NanoVDB, tables) {setup_s:.1f} s", file=sys.stderr, flush=True)
    net = kpn.KpnDenoiser(seed=0)
    gen = torch.Generator(device="cpu").manual_seed(0)
    pairs = []
    cs = C5["crop"]

    def frame(k, log):
        ts = {}
        _c5_camera(vpt, k)
        vpt.set_use_feature_maps([])
        vpt.set_seed_offset(2 * k)
        t = time.perf_counter()
        img = vpt.render_frame(target, spp).clone()
        torch.cuda.synchronize()
        ts["render_s"] = time.perf_counter() - t
        ts["path_kernel_ms"] = vpt.get_last_render_stats()["path_kernel_ms"]
        t = time.perf_counter()
        vpt.set_seed_offset(2 * k + 1)
        f, x = kpn.render_inputs(vpt, H, W, pspp)
        pairs.append((f, x, img))
        torch.cuda.synchronize()
        ts["pair_s"] = time.perf_counter() - t
        t = time.perf_counter()
        for _ in range(C5["train_steps"]):
            fs, xs, ys = [], [], []
            for _ in range(C5["batch"]):
                pf, px, py = pairs[int(torch.randint(0, len(pairs), (1,), generator=gen))]
                oy = int(torch.randint(0, H - cs + 1, (1,), generator=gen))
                ox = int(torch.randint(0, W - cs + 1, (1,), generator=gen))
                fs.append(pf[:, oy:oy + cs, ox:ox + cs])
                xs.append(px[:, oy:oy + cs, ox:ox + cs])
                ys.append(py[:, oy:oy + cs, ox:ox + cs])
            loss = net.train_step(torch.stack(fs), torch.stack(xs), torch.stack(ys), C5["lr"])
        torch.cuda.synchronize()
        ts["train_s"] = time.perf_counter() - t
        t = time.perf_counter()
        den = net.denoise(f.unsqueeze(0), x.unsqueeze(0))[0]
        torch.cuda.synchronize()
        ts["inference_s"] = time.perf_counter() - t
        ts["train_loss"] = loss
        ts["noisy_rel_mse"] = kpn.rel_mse(x, img)
        ts["denoised_rel_mse"] = kpn.rel_mse(den, img)
        print(f"{log} {k}: " + ", ".join(f"{a} {b:.4g}" for a, b in ts.items()), file=sys.stderr, flush=True)
        return ts

    for k in range(args.warmup):
        frame(k, "warmup frame")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    per = [frame(args.warmup + k, "frame") for k in range(args.steps)]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    paths = W * H * (spp + pspp)
    value = paths / (wall / args.steps) / 1e6
    # instrumented pass of the production render (untimed): device work counters for the algorithmic bytes
    vpt.set_use_feature_maps([])
    vpt.set_collect_stats(True)
    vpt.set_seed_offset(0)
    _c5_camera(vpt, 0)
    vpt.render_frame(target, spp)
    torch.cuda.synchronize()
    c = vpt.get_last_render_counters()
    vpt.set_collect_stats(False)
    kernel_ms = float(np.mean([p["path_kernel_ms"] for p in per]))
    launches = vpt.get_last_render_stats()["launches"]
    # Fp16 leaves: 2-byte codes per tap; a leaf-block lookup reads the 4-byte block-index entry and the leaf's
    # minimum and quantum (8 bytes). The frame's launches trace equal path counts, so per-launch figures are the
    # frame's divided by the launch count.
    b_alg = 2 * c["density_taps"] + 8 * c["supervoxel_reads"] + 12 * c["tree_descents"] + 16 * c["paths"]
    achieved = b_alg / (kernel_ms * 1e-3) / 1e9
    keys = ("render_s", "pair_s", "train_s", "inference_s")
    line = {
        "metric": "Mpaths/s on 1 MI355X", "value": round(value, 3), "unit": "Mpaths/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded fBm procedural cloud generated on the GPU)",
        "config": {"workload": "c5_fp16_2048", "volume": f"NanoVDB Fp16 leaves, 2048^3 anvil fBm seed 0, "
                   f"smoothstep({C5['edge0']}, {C5['edge1']})", "image": f"{W}x{H}", "spp_per_frame": spp,
                   "pair_spp": pspp, "estimator": "residual ratio tracking", "phase_g": 0.5, "albedo": 0.9,
                   "extinction": 1024.0, "max_bounces": C5["max_bounces"], "majorant": "16^3 supervoxel grid",
                   "kpn_train_steps_per_frame": C5["train_steps"], "kpn_batch": C5["batch"], "crop": cs,
                   "paths_per_step": paths},
        "setup_s": round(setup_s, 2), "nanovdb": nano,
        "per_frame_s": {k: round(float(np.mean([p[k] for p in per])), 4) for k in keys},
        "rel_mse_vs_frame": {"noisy_8spp": [round(p["noisy_rel_mse"], 6) for p in per],
                             "denoised_8spp": [round(p["denoised_rel_mse"], 6) for p in per]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": load_traffic("c5_fp16_2048"),
                     "kernel": "k_pathPool<RESIDUAL_RATIO_TRACKING, NanoAccessT<false, false>, false>", "launches_per_frame": launches,
                     "launch_ms": round(kernel_ms / launches, 3), "algorithmic_bytes_per_launch": int(b_alg / launches),
                     "counters_per_frame": {k: int(v) for k, v in c.items()}},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def host_threads():
    """The host threads this process may use: OMP_NUM_THREADS when set (on a shared GPU host it is the CPU share of
    one GPU), else every CPU in the process's affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def host_cpu():
    """Host CPU model (lscpu's "Model name", read from /proc/cpuinfo) and the logical CPU count."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "threads_used": host_threads()}


def cpu_baseline_c1(budget_s: float = 12.0):
    """The oracle (FP32 restatement of the reference shaders, hybrid-Taus RNG) on the C1 scene and camera."""
    import numpy as np

    from oracle import oracle as O

    threads = host_threads()
    d = O.fbm_cloud(C1["grid"], C1["grid"], C1["grid"], seed=0)
    sun = np.array([0.5826, 0.7660, 0.2717])
    sun = 100.0 * sun / np.linalg.norm(sun)
    L = O.make_light(light_type=2, light_space=0, position=tuple(sun), intensity=2.6,
                     color=(1.0, 0.961538462, 0.884615385))
    res = 64
    sc = O.Scene(d, mode="next_event", interpolation="stochastic", extinction=1024.0, albedo=0.9, g=0.85,
                 env_image=np.ones((3, 16, 32), np.float32), lights=[L], width=res, height=res, max_grid_extent=0.5,
                 max_bounces=C1["max_bounces"])
    return _time_oracle(sc, res, threads, budget_s, "C1 scene and camera")


def cpu_baseline_c2(vpt, budget_s: float = 15.0):
    """The oracle on the C2 volume (its dense equivalent: the oracle samples the same normalised voxels) and camera,
    cycling through the same estimator / g combinations, on a bounded sample of 64 x 64 pixels."""
    import numpy as np

    from oracle import oracle as O

    threads = host_threads()
    g = C2["grid"]
    d = vpt.get_density_field((g, g, g), normalized=True)
    # The sparse grid spans the active voxels only: crop the dense field to them and use the same world box.
    nzm = d != 0
    zs, ys, xs = (np.nonzero(nzm.any(axis=ax))[0] for ax in ((1, 2), (0, 2), (0, 1)))
    d = np.ascontiguousarray(d[zs[0]:zs[-1] + 1, ys[0]:ys[-1] + 1, xs[0]:xs[-1] + 1])
    del nzm
    box = vpt.get_render_bounding_box()
    box = (np.array(box[0::2], np.float32), np.array(box[1::2], np.float32))
    res = 64
    names = {0: "delta", 2: "ratio", 4: "residual_ratio", 3: "decomposition"}
    paths, elapsed, frames = 0, 0.0, 0
    combos = [(m, gg) for m, _ in C2["modes"] for gg in C2["gs"]]
    scenes = {}
    while elapsed < budget_s and frames < 4 * len(combos):
        m, gg = combos[frames % len(combos)]
        if (m, gg) not in scenes:
            scenes[(m, gg)] = O.Scene(d, mode=names[m], interpolation="stochastic", extinction=1024.0, albedo=0.9,
                                      g=gg, width=res, height=res, voxel_range=(0.0, 1.0), box=box,
                                      super_voxel_size=16, max_bounces=C2["max_bounces"])
        t0 = time.perf_counter()
        scenes[(m, gg)].render(1, frame_start=frames, threads=threads)
        elapsed += time.perf_counter() - t0
        paths += res * res
        frames += 1
    return {"value": paths / elapsed / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"{res}x{res} px x {frames} spp ({paths} paths) of the C2 volume and camera, estimator/g "
                      f"cycling like the timed steps, {elapsed:.1f} s on {threads} threads",
            "cpu": host_cpu()}


def _time_oracle(sc, res, threads, budget_s, what):
    paths, elapsed, frames = 0, 0.0, 0
    acc = None
    while elapsed < budget_s and frames < 64:
        t0 = time.perf_counter()
        acc, _ = sc.render(1, frame_start=frames, acc=acc, threads=threads)
        elapsed += time.perf_counter() - t0
        paths += res * res
        frames += 1
    return {"value": paths / elapsed / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"{res}x{res} px x {frames} spp ({paths} paths) of the {what}, {elapsed:.1f} s on "
                      f"{threads} threads", "cpu": host_cpu()}


def load_pmc(workload: str) -> dict:
    """The dominant kernel's committed rocprofv3 PMC summary (profiles/pmc_<workload>.json), or {}."""
    path = os.path.join(REPO, "profiles", f"pmc_{workload}.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def load_traffic(workload: str):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 PMC summary, when one exists."""
    return load_pmc(workload).get("traffic_bytes_per_launch")


def valu_issue(workload: str):
    """The dominant kernel's vector-instruction issue from the committed PMC summary: VALU instructions per SIMD per
    cycle (SQ_INSTS_VALU over 1024 SIMDs and the per-XCD GRBM_GUI_ACTIVE cycles) and the fraction of lanes active per
    issued instruction (SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU)), or None without one."""
    c = load_pmc(workload).get("counters_per_dispatch", {})
    if not all(k in c for k in ("SQ_INSTS_VALU", "GRBM_GUI_ACTIVE", "SQ_THREAD_CYCLES_VALU", "SQ_ACTIVE_INST_VALU")):
        return None
    return {"insts_per_simd_cycle": round(c["SQ_INSTS_VALU"] / 1024 / (c["GRBM_GUI_ACTIVE"] / 8), 4),
            "lane_utilisation": round(c["SQ_THREAD_CYCLES_VALU"] / (64 * c["SQ_ACTIVE_INST_VALU"]), 4)}


def copy_calibration(torch):
    """A plain streaming copy kernel of 256 MiB read + 256 MiB written (an elementwise multiply by 1, so it runs as
    a kernel dispatch the PMC counters see): the FETCH_SIZE / WRITE_SIZE scale factor for the PMC summaries."""
    n = 256 * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}), flush=True)


def kpn_inference(torch, calls: int, res: int = 1024):
    """KPN inference alone (packing, six bf16 MFMA convolutions, softmax filter) on seeded inputs of one res x res
    frame, `calls` times: the dispatches the MFMA counter passes profile."""
    from cloudrendering_amd import kpn, vpt

    vpt.initialize()
    net = kpn.KpnDenoiser(seed=1)
    g = torch.Generator(device="cuda").manual_seed(0)
    f = torch.rand((1, kpn.NUM_INPUTS, res, res), device="cuda", generator=g)
    x = torch.rand((1, 3, res, res), device="cuda", generator=g)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(calls):
        y = net.denoise(f, x)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) * 1e3 / calls
    print(json.dumps({"kpn_inference_calls": calls, "res": res, "ms_per_call": round(ms, 4),
                      "output_mean": float(y.mean())}), flush=True)
    vpt.cleanup()


def denoise_frames(torch, frames: int, res: int = 1024):
    """EAW (C4's 4 passes on 4-spp frames) and SVGF (1-spp frames) on the denoiser workload's cloud at res x res:
    the dispatches the denoiser kernel traces profile. Prints the wall time per frame of each."""
    import numpy as np

    from cloudrendering_amd import vpt

    vpt.initialize()
    _c4_scene(vpt, 999)
    _c4_camera(vpt, np.random.default_rng(999))
    probe = torch.zeros((3, res, res), dtype=torch.float32, device="cuda")
    out = {"res": res, "frames": frames}
    for name, spp, prop in (("EAW", C4["input_spp"], str(C4["eaw_iterations"])), ("SVGF", 1, "5")):
        vpt.set_denoiser(name)
        vpt.set_denoiser_property("maxNumIterations", prop)
        vpt.render_frame(probe, spp)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(frames):
            vpt.render_frame(probe, spp)
        torch.cuda.synchronize()
        out[f"{name.lower()}_ms_per_frame"] = round((time.perf_counter() - t) * 1e3 / frames, 3)
    vpt.set_denoiser("None")
    print(json.dumps(out), flush=True)
    vpt.cleanup()


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0,
                    help="timed steps (0: 2 sweeps of the 12 estimator x g renders for c2_sparse512, 12 for "
                         "c1_dense128, 3 otherwise)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["c2_sparse512", "c1_dense128", "c3_batch", "c4_kpn", "c5_fp16_2048"],
                    default="c2_sparse512")
    ap.add_argument("--kpn-crops", type=int, default=0)
    ap.add_argument("--kpn-steps", type=int, default=0)
    ap.add_argument("--kpn-eval-res", type=int, default=0)
    ap.add_argument("--kpn-ref-spp", type=int, default=0)
    ap.add_argument("--spp", type=int, default=0, help="samples per pixel per step (0: the workload's)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kpn", action="store_true",
                    help="c2_sparse512: skip the denoiser half of the metric (the C4 workload after the timed steps)")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--kpn-inference", type=int, default=0, metavar="CALLS",
                    help="profiling helper: CALLS KPN inferences on seeded 1024^2 inputs, then exit")
    ap.add_argument("--denoise-frames", type=int, default=0, metavar="FRAMES",
                    help="profiling helper: FRAMES 4-spp EAW (4 passes) and FRAMES 1-spp SVGF renders of the "
                         "denoiser workload's cloud at 1024^2, then exit")
    args = ap.parse_args()
    if args.steps <= 0:
        args.steps = {"c2_sparse512": 2, "c1_dense128": 12}.get(args.workload, 3)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path tracer is single-GPU here",
              file=sys.stderr)
        return 2

    import torch

    if args.copy_calibration:
        copy_calibration(torch)
        return 0
    if args.kpn_inference > 0:
        kpn_inference(torch, args.kpn_inference)
        return 0
    if args.denoise_frames > 0:
        denoise_frames(torch, args.denoise_frames)
        return 0

    from cloudrendering_amd import vpt

    vpt.initialize()
    if args.workload in ("c3_batch", "c4_kpn", "c5_fp16_2048"):
        {"c3_batch": run_c3, "c4_kpn": run_c4, "c5_fp16_2048": run_c5}[args.workload](args, vpt, torch)
        vpt.cleanup()
        return 0
    dev = torch.device("cuda:0")
    c2 = args.workload == "c2_sparse512"
    if c2:
        setup_c2(vpt)
        W, H, spp = C2["width"], C2["height"], args.spp or C2["spp"]
        combos = [(m, g) for m, _ in C2["modes"] for g in C2["gs"]]
    else:
        setup_c1(vpt, torch)
        W, H, spp = C1["width"], C1["height"], args.spp or C1["spp"]
        combos = [(5, 0.85)]
    target = torch.zeros((3, H, W), dtype=torch.float32, device=dev)

    def apply(k):
        m, g = combos[k % len(combos)]
        vpt.set_vpt_mode(m)
        vpt.set_phase_g(g)
        return m

    # One step renders every combination once (c2_sparse512: the 12 estimator x g renders; c1_dense128: one).
    for k in range(args.warmup):
        for c in range(len(combos)):
            apply(c)
            vpt.render_frame(target, spp)
    torch.cuda.synchronize()
    per_combo_ms = {}
    t0 = time.perf_counter()
    for k in range(args.steps):
        for c in range(len(combos)):
            apply(c)
            vpt.render_frame(target, spp)
            st = vpt.get_last_render_stats()
            per_combo_ms.setdefault(combos[c], []).append(st["path_kernel_ms"])
            launches = st["launches"]
            print(f"step {k}: {combos[c]} path kernel {st['path_kernel_ms']:.1f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    paths_per_step = W * H * spp * len(combos)
    value = paths_per_step / (wall / args.steps) / 1e6

    # Dominant kernel: the estimator whose path kernel takes the most time over the timed steps.
    by_mode = {}
    for (m, g), ms in per_combo_ms.items():
        by_mode.setdefault(m, []).extend(ms)
    dom = max(by_mode, key=lambda m: sum(by_mode[m]))
    dom_ms = sum(by_mode[dom]) / len(by_mode[dom])
    # Instrumented passes (untimed) of the dominant estimator's combinations: device work counters over the same
    # image and sample count (the counters add wave-reduced atomics, so these passes are not timed).
    taps = svr = paths = desc = 0
    n_c = 0
    vpt.set_collect_stats(True)
    for k in range(len(combos)):
        if combos[k][0] != dom or combos[k] not in per_combo_ms:
            continue
        apply(k)
        vpt.render_frame(target, spp)
        torch.cuda.synchronize()
        c = vpt.get_last_render_counters()
        taps += c["density_taps"]
        svr += c["supervoxel_reads"]
        paths += c["paths"]
        desc += c["tree_descents"]
        n_c += 1
    vpt.set_collect_stats(False)
    taps, svr, paths, desc = (v / max(n_c, 1) for v in (taps, svr, paths, desc))
    # Algorithmic bytes per render: 4-byte FP32 density taps, 4-byte majorant / 8-byte (control, residual) table
    # reads, 8 bytes per leaf-block lookup on a leaf-cache miss (the 4-byte entry of the flat 8^3-block index and the
    # 4-byte tile value of a block without a leaf), and the 16-byte radiance record of each path.
    b_alg = 4 * taps + (4 if dom in (0, 2, 5) else 8) * svr + 8 * desc + 16 * paths
    achieved = b_alg / (dom_ms * 1e-3) / 1e9
    if c2:
        kernel = MODE_KERNELS[dom]
        config = {"workload": args.workload, "volume": "NanoVDB fp32 512^3 fBm seed 0, ~12% active, 8^3 leaves",
                  "image": f"{W}x{H}", "spp_per_step": spp,
                  "estimators": [n for _, n in C2["modes"]], "phase_g": C2["gs"], "albedo": 0.9,
                  "extinction": 1024.0, "environment": "default sky", "max_bounces": C2["max_bounces"],
                  "majorant": "16^3 supervoxel grid",
                  "step": "one render of each of the 12 estimator x g combinations (12 x 268,435,456 paths)"}
    else:
        kernel = "k_pathPool<NEXT_EVENT_TRACKING, DenseAccessT<false>, false>"
        config = {"workload": args.workload, "volume": "dense 128^3 fp32 fBm seed 0", "image": f"{W}x{H}",
                  "spp_per_step": spp, "estimator": "next event tracking", "phase_g": 0.85, "albedo": 0.9,
                  "extinction": 1024.0, "environment": "constant white image + directional sun",
                  "max_bounces": C1["max_bounces"], "majorant": "8^3 supervoxel grid"}
    per_combo = {f"{m}/{g}": round(sum(v) / len(v), 3) for (m, g), v in per_combo_ms.items()}
    names = dict(C2["modes"]) if c2 else {5: "next event tracking"}
    # path-kernel rate of each estimator over its timed renders (all g)
    per_mode = {names[m]: {"path_kernel_ms": round(sum(ms) / len(ms), 3),
                           "mpaths_per_s": round(W * H * spp / (sum(ms) / len(ms) * 1e-3) / 1e6, 2)}
                for m, ms in by_mode.items()}
    line = {
        "metric": "Mpaths/s on 1 MI355X",
        "value": round(value, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded fBm procedural cloud generated on the GPU)",
        "config": config,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": load_traffic(args.workload),
                     "valu_insts_per_launch": load_pmc(args.workload).get("valu_insts_per_launch"),
                     "valu_issue": valu_issue(args.workload),
                     "kernel": kernel, "launches_per_step": launches,
                     "launch_ms": round(dom_ms / launches, 4), "algorithmic_bytes_per_launch": int(b_alg / launches),
                     "density_taps": int(taps), "supervoxel_reads": int(svr), "tree_descents": int(desc),
                     "paths": int(paths)},
        "kernel_ms_per_mode_g": per_combo,
        "per_estimator": per_mode,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline_c2(vpt) if c2 else cpu_baseline_c1()
    if c2 and not args.no_kpn:
        # The second half of the metric, measured in the same run (after the timed steps): the C4 denoiser workload's
        # wall time to relMSE 1e-3 against plain tracking, with its equal-time comparison (see run_c4).
        k = run_c4(args, vpt, torch, emit=False)
        line["kpn_wall_time_to_rel_mse_1e-3"] = {
            "value_s": k["value"], "plain_s": k["plain_seconds_to_1e-3"], "speedup": k["speedup_to_1e-3"],
            "interpolated_s": k["kpn_seconds_to_1e-3_interpolated"],
            "unit": "s", "higher_is_better": False, "config": k["config"], "kpn": k["kpn"], "plain": k["plain"],
            "equal_time_4spp": k["equal_time_4spp"], "inference": k["inference"],
            "pair_generation_s": k["pair_generation_s"], "train_s_per_network": k["train_s_per_network"]}
    vpt.cleanup()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
