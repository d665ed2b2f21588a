"""Throughput benchmark of the gfx950 volumetric path tracer on its headline workload.

Workload (one GPU): dense 128^3 FP32 seeded fBm cloud (5 octaves, seed 0, lacunarity 2, gain 0.5, base frequency
4, smoothstep(0.45, 0.6)) filling the unit cube, extinction 1024, albedo 0.9, Henyey-Greenstein g = 0.85, constant
white environment image plus one directional sun with next-event estimation, 256 x 256 pinhole camera at the
reference module's default pose, 64 samples per pixel per step, unbounded path length (the reference trackers have
no bounce limit). A step is one render_frame pass of 256 * 256 * 64 camera paths with the volume resident in HBM.

Prints one JSON line: Mpaths/s, the roofline of the dominant kernel (algorithmic bytes from device work counters of
an instrumented pass with identical seeds / HIP-event kernel time of the timed passes), and the CPU oracle timed
on a bounded sample of the same scene on the host cores.
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
W = H = 256
SPP = 64
GRID = 128


def setup_scene(vpt, torch):
    vpt.set_max_grid_extent(0.5)
    vpt.generate_fbm_cloud(GRID, GRID, GRID, seed=0, octaves=5, lacunarity=2.0, gain=0.5, base_frequency=4.0,
                           edge0=0.45, edge1=0.6)
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


def cpu_baseline(budget_s: float = 12.0):
    """The oracle (FP32 restatement of the reference shaders, hybrid-Taus RNG) on the same scene and camera, timed on
    the host cores on a bounded sample: a 64 x 64 image of the same view at growing spp until the budget is used."""
    import numpy as np

    from oracle import oracle as O

    threads = max(1, min(os.cpu_count() or 1, 16))
    d = O.fbm_cloud(GRID, GRID, GRID, seed=0)
    sun = np.array([0.5826, 0.7660, 0.2717])
    sun = 100.0 * sun / np.linalg.norm(sun)
    L = O.make_light(light_type=2, light_space=0, position=tuple(sun), intensity=2.6,
                     color=(1.0, 0.961538462, 0.884615385))
    res = 64
    sc = O.Scene(d, mode="next_event", interpolation="stochastic", extinction=1024.0, albedo=0.9, g=0.85,
                 env_image=np.ones((3, 16, 32), np.float32), lights=[L], width=res, height=res, max_grid_extent=0.5)
    paths, elapsed, frames = 0, 0.0, 0
    acc = None
    while elapsed < budget_s and frames < 64:
        t0 = time.perf_counter()
        acc, _ = sc.render(1, frame_start=frames, acc=acc, threads=threads)
        elapsed += time.perf_counter() - t0
        paths += res * res
        frames += 1
    return {"value": paths / elapsed / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"{res}x{res} px x {frames} spp ({paths} paths) of the same scene and camera, "
                      f"{elapsed:.1f} s on {threads} threads"}


def load_traffic(workload: str):
    """HBM bytes per launch of the path kernel from the committed rocprofv3 PMC summary, when one exists."""
    path = os.path.join(REPO, "profiles", f"pmc_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        pmc = json.load(f)
    return pmc.get("traffic_bytes_per_launch")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path tracer is single-GPU here",
              file=sys.stderr)
        return 2

    import torch

    from cloudrendering_amd import vpt

    vpt.initialize()
    setup_scene(vpt, torch)
    dev = torch.device("cuda:0")
    target = torch.zeros((3, H, W), dtype=torch.float32, device=dev)

    # Instrumented pass (untimed): algorithmic work counters of one step.
    vpt.set_collect_stats(True)
    vpt.render_frame(target, SPP)
    torch.cuda.synchronize()
    counters = vpt.get_last_render_counters()
    vpt.set_collect_stats(False)

    for _ in range(args.warmup):
        vpt.render_frame(target, SPP)
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        vpt.render_frame(target, SPP)
        kernel_ms.append(vpt.get_last_render_stats()["path_kernel_ms"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    paths_per_step = W * H * SPP
    value = paths_per_step / (wall / args.steps) / 1e6

    # Algorithmic bytes per launch of the path kernel: FP32 density taps, 4-byte majorant reads, 16-byte path record.
    b_alg = 4 * counters["density_taps"] + 4 * counters["supervoxel_reads"] + 16 * counters["paths"]
    k_ms = sum(kernel_ms) / len(kernel_ms)
    achieved = b_alg / (k_ms * 1e-3) / 1e9
    workload = "c1_dense128_nee_256x256_64spp"
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
        "config": {"workload": workload, "volume": "dense 128^3 fp32 fBm seed 0", "image": f"{W}x{H}",
                   "spp_per_step": SPP, "estimator": "next event tracking", "phase_g": 0.85, "albedo": 0.9,
                   "extinction": 1024.0, "environment": "constant white image + directional sun",
                   "max_bounces": "unbounded", "majorant": "8^3 supervoxel grid"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": load_traffic(workload),
                     "kernel": "k_pathPool<NEXT_EVENT_TRACKING, DenseAccess>", "kernel_ms": round(k_ms, 3),
                     "algorithmic_bytes": b_alg, "density_taps": counters["density_taps"],
                     "supervoxel_reads": counters["supervoxel_reads"], "paths": counters["paths"]},
    }
    vpt.cleanup()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
