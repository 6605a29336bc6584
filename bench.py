#!/usr/bin/env python3
"""Benchmark of the polygon path on one MI355X.

Workload: the 1920x1080 textured terrain frame -- a 256x256-point grid (130,050 triangles),
a 1024x1024 RGBA8 texture with 11 mip levels, per-vertex Lambert colours, a 60 degree camera
30 units up pitched down 35 degrees (dfpsr_amd/scenes.py: terrain()).  Synthetic, seeded.

One step = one frame through the C ABI, exactly as a reference application draws one:
clear both canvases (image_fill), renderer_begin, renderer_giveTask(model), renderer_end.
Model, texture and canvases stay resident in HBM; renderer_end returns with the frame drawn.

Reported: shaded Mpixel/s (canvas pixels per second), Mtriangle/s (submitted triangles per
second), frames/s, the whole frame's algorithmic GB/s, the roofline of the dominant kernel
(the tile raster kernel, timed live with HIP events on the renderer's stream) and the
reference CPU renderer timed on this machine's host cores.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WIDTH, HEIGHT = 1920, 1080
POINTS = 256 * 256
TRIANGLES = 130050
TEXELS_L0_TO_L4 = sum((1024 >> l) * (1024 >> l) for l in range(5))  # sampling never goes past level 4
# Algorithmic bytes per frame (colour write + depth read + depth write per canvas pixel,
# sampled texture levels, positions, per-triangle index + colour + uv).
FRAME_BYTES = 12 * WIDTH * HEIGHT + 4 * TEXELS_L0_TO_L4 + 12 * POINTS + 84 * TRIANGLES
# The raster kernel's share: the canvas traffic and the sampled texels.
RASTER_BYTES = 12 * WIDTH * HEIGHT + 4 * TEXELS_L0_TO_L4


def cpu_baseline(iterations: int = 20, warmup: int = 3) -> dict:
    """Times the reference renderer (compiled from its sources into oracle/_ref) on this host;
    falls back to the oracle's restatement when that build is absent."""
    from oracle import harness
    kind = "reference" if harness.ref_available() else "port"
    if kind == "reference":
        harness.load_ref()  # before large numpy work
    from dfpsr_amd import scenes
    sc = scenes.terrain()
    nproc = os.cpu_count() or 1
    threads = min(max(nproc - 1, 1), 12)  # threading.cpp:70-72 with CommandQueue's 12 jobs
    if kind == "reference":
        seconds = harness.ref_time(sc, warmup, iterations)
    else:
        seconds = harness.oracle_time(sc, threads, warmup, iterations)
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(WIDTH * HEIGHT / seconds / 1e6, 3), "unit": "Mpix/s", "cores": threads, "kind": kind,
            "sample": f"median of {iterations} frames (after {warmup} warm-up) of the 1920x1080 terrain, "
                      f"renderer_begin..renderer_end, {threads} raster threads of {nproc} ({cpu_model})",
            "ms_per_frame": round(seconds * 1e3, 3), "mtri_per_s": round(TRIANGLES / seconds / 1e6, 3)}


def traffic_from_profiles():
    """HBM bytes per raster launch measured with rocprofv3 --pmc (profiles/), if recorded."""
    path = os.path.join(ROOT, "profiles", "raster_terrain_1080p_traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get("traffic_bytes")
    except (OSError, ValueError):
        return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-baseline-only", action="store_true")
    args = ap.parse_args()
    if args.cpu_baseline_only:
        print(json.dumps(cpu_baseline()))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(every box here has a single MI355X).", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from dfpsr_amd import api, scenes

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    sc = scenes.terrain()
    gpu_scene = api.GpuScene(sc)
    renderer = api.renderer_create()
    color = torch.zeros((HEIGHT, WIDTH), dtype=torch.int32, device="cuda")
    depth = torch.zeros((HEIGHT, WIDTH), dtype=torch.float32, device="cuda")
    color_im, depth_im = api.device_image(color), api.device_image(depth)

    def step():
        color.zero_()
        depth.zero_()
        api.renderer_begin(renderer, color_im, depth_im)
        gpu_scene.give(renderer)
        api.renderer_end(renderer)

    for _ in range(args.warmup):
        step()
    renderer.set_timing(True)
    raster_ms = []
    stage_ms = np.zeros(5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        t = renderer.timing()
        raster_ms.append(float(t[3]))
        stage_ms += t
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    stats = renderer.stats()

    ms_per_step = elapsed / args.steps * 1e3
    raster_avg_ms = float(np.mean(raster_ms))
    achieved = RASTER_BYTES / (raster_avg_ms * 1e-3) / 1e9
    line = {
        "metric": "shaded Mpix/s at 1080p (textured terrain, 130050 triangles)",
        "value": round(WIDTH * HEIGHT * args.steps / elapsed / 1e6, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+u8",
        "data": "synthetic (seeded terrain, noise+checker texture)",
        "config": {"workload": "1920x1080 textured terrain frame, 256x256 grid, 130050 triangles, 1024^2 RGBA8 "
                               "texture with 11 mips, renderer_begin/giveTask/end with canvas clears",
                   "width": WIDTH, "height": HEIGHT, "triangles": TRIANGLES, "parallelism": "single GPU"},
        "mtri_per_s": round(TRIANGLES * args.steps / elapsed / 1e6, 3),
        "frames_per_s": round(args.steps / elapsed, 2),
        "frame_gbs": round(FRAME_BYTES / (ms_per_step * 1e-3) / 1e9, 2),
        "stage_ms": {k: round(float(v) / args.steps, 4) for k, v in
                     zip(["project", "setup", "binning", "raster", "device_frame"], stage_ms)},
        "draw_commands": int(stats[0]),
        "tile_pairs": int(stats[1]),
        "roofline": {"bound": "hbm", "kernel": "k_raster", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(),
                     "algorithmic_bytes": RASTER_BYTES, "avg_launch_ms": round(raster_avg_ms, 5)},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        try:
            res = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-only"], capture_output=True,
                                 text=True, timeout=300)
            lines = [l for l in res.stdout.splitlines() if l.startswith("{")]
            if lines:
                line["cpu_baseline"] = json.loads(lines[-1])
        except (subprocess.TimeoutExpired, ValueError):
            pass
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
