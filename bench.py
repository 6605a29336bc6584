#!/usr/bin/env python3
"""Benchmark of the polygon path on one MI355X.

Workload: the 1920x1080 textured terrain frame -- a 256x256-point grid (130,050 triangles),
a 1024x1024 RGBA8 texture with 11 mip levels, per-vertex Lambert colours, a 60 degree camera
30 units up pitched down 35 degrees (dfpsr_amd/scenes.py: terrain()).  Synthetic, seeded.

One step = one frame through the C ABI, as a reference application draws one: clear both
canvases (image_fill of colour 0 and depth 0, done by the renderer's clear option inside the
frame), renderer_begin, renderer_giveTask(model), renderer_end.  Model, texture and canvases
stay resident in HBM; renderer_end returns with the frame drawn.  The renderer's tuning (tile
shape, hipGraph replay) is the fastest of a short untimed sweep unless given by flags.

Reported: shaded Mpixel/s (canvas pixels per second), Mtriangle/s (submitted triangles per
second), frames/s, the whole frame's algorithmic GB/s, the roofline of the dominant kernel
(the frame's 42,180,776 algorithmic bytes over the average launch of the tile raster kernel,
timed live with HIP events on the renderer's stream, beside the same bytes over the whole
device frame and the raster's own share of them) and the reference CPU renderer timed on this
machine's host cores.

Two more objects ride on the same line (skip them with --no-batch / --no-diff):
  "batched": 4,096 uneven jobs (canvases 64..640 x 48..480, 100..20,000 triangles, cube soups
      and terrain patches, shared 256^2 textures) drawn by one batched, graph-replayed launch
      sequence with the autotuned tile shape, against the per-job loop of the plain renderer
      on the same GPU; images compared for identity.
  "differentiable": forward + adjoint steps of the 8-view 1280x720 fit (50,000-triangle
      sphere, 512^2 float texture, per-point RGB multipliers) and of the 256-view 256x256
      batch (20,480 triangles, 1024^2 texture).
  "depth_only": model_renderDepth of the 1920x1080 terrain into a device depth canvas (skip
      with --no-depth), per synchronous call.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--no-batch] [--no-diff]
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


def progress(message: str) -> None:
    print(f"[bench] {message}", file=sys.stderr, flush=True)


def batched_form(repeats: int = 5) -> dict:
    """4,096 uneven jobs: the batched form (autotuned, graph replay) against the per-job loop of
    the plain renderer, images compared for identity."""
    import torch
    from dfpsr_amd import api, autotune, scenes
    t0 = time.perf_counter()
    jobs = scenes.batch_jobs()
    gpu = api.GpuJobs(jobs)
    setup_s = time.perf_counter() - t0
    progress(f"batched: {len(jobs)} jobs resident after {setup_s:.1f} s")
    batch = api.Batch()
    stored = autotune.load()
    if stored is None:
        best, table = batch.autotune(gpu.jobs, 3, api.tuning(clear=True))
        tile_config, per_cu, source = int(best.tile_config), int(best.persistent_per_cu), "swept in this run"
    else:
        (tile_config, per_cu), source = stored, "dfpsr_amd/tuning_gfx950.json"
    t = api.tuning(tile_config=tile_config, persistent_per_cu=per_cu, use_graph=True, clear=True)
    for _ in range(2):
        batch.render(gpu.jobs, t)
    torch.cuda.synchronize()
    start = time.perf_counter()
    for _ in range(repeats):
        batch.render(gpu.jobs, t)
    torch.cuda.synchronize()
    batch_s = (time.perf_counter() - start) / repeats
    progress(f"batched: {batch_s * 1e3:.2f} ms per batch")
    batched_colors = [c.clone() for c in gpu.colors]
    batched_depths = [d.clone() for d in gpu.depths]
    renderer = api.renderer_create()
    api.renderer_setTuning(renderer, api.tuning(clear=True))
    gpu.render_plain(renderer)  # warm-up pass
    torch.cuda.synchronize()
    start = time.perf_counter()
    gpu.render_plain(renderer)
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - start
    progress(f"batched: plain loop {plain_s * 1e3:.2f} ms")
    identical = all(torch.equal(a, b) for a, b in zip(batched_colors, gpu.colors)) and \
        all(torch.equal(a.view(torch.int32), b.view(torch.int32)) for a, b in zip(batched_depths, gpu.depths))
    stats = batch.stats()
    n = len(jobs)
    return {"workload": f"{n} uneven jobs (seed 2): {int(sum(j.triangle_count for j in jobs))} triangles, "
                        f"{int(sum(j.width * j.height for j in jobs))} canvas pixels",
            "frames_per_s": round(n / batch_s, 1), "ms_per_batch": round(batch_s * 1e3, 3),
            "plain_loop_frames_per_s": round(n / plain_s, 1), "plain_loop_ms": round(plain_s * 1e3, 3),
            "speedup_vs_plain_loop": round(plain_s / batch_s, 2), "identical_images": bool(identical),
            "tuning": {"tile": autotune.TILE_SHAPES[tile_config], "persistent_per_cu": per_cu, "use_graph": 1,
                       "source": source},
            "draw_commands": int(stats[0]), "tile_pairs": int(stats[1]), "setup_s": round(setup_s, 1)}


def differentiable_form(steps: int = 10) -> dict:
    """Forward + adjoint steps of the two differentiable workloads."""
    import torch
    from dfpsr_amd import api, scenes
    out = {}
    for key, views, tex_size, with_mul in (("fit_8x1280x720", scenes.sphere_views(50, 8, 1280, 720, seed=3), 512, True),
                                           ("batch_256x256x256", scenes.sphere_views(32, 256, 256, 256, seed=11), 1024, False)):
        gv = api.GpuViews(views)
        tex = torch.from_numpy(scenes.float_texture(tex_size, 5, smooth=16)).cuda()
        mul = torch.ones((gv.point_count, 3), dtype=torch.float32, device="cuda") if with_mul else None
        gt = torch.zeros_like(tex)
        gm = torch.zeros_like(mul) if with_mul else None
        for g in gv.grad:
            g.fill_(1e-3)
        r = api.DiffRenderer()

        def timed(fn, count):
            torch.cuda.synchronize()
            start = time.perf_counter()
            for _ in range(count):
                fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - start) / count

        # Untimed choice of the tile shape, and for each pass of the persistent workgroups per CU
        # (0 = one per tile): per tile shape the forward's best setting, then the adjoint's on that
        # forward's output; the shape with the shortest step wins.
        best = None
        for tile_config in (0, 1, 2):
            choice = {}
            total = 0.0
            for stage in ("forward", "backward"):
                stage_best = None
                for k in (0, 4):
                    tk = api.tuning(clear=True, tile_config=tile_config, persistent_per_cu=k, use_graph=True)
                    if stage == "forward":
                        r.forward(gv.model, gv.views, tex, mul, tk)
                        ms = timed(lambda: r.forward(gv.model, gv.views, tex, mul, tk), 3)
                    else:
                        r.backward(gt, gm, tk)
                        ms = timed(lambda: r.backward(gt, gm, tk), 3)
                    if stage_best is None or ms < stage_best[0]:
                        stage_best = (ms, k)
                choice[stage] = stage_best[1]
                total += stage_best[0]
                if stage == "forward":  # the adjoint reads this forward's output
                    r.forward(gv.model, gv.views, tex, mul,
                              api.tuning(clear=True, tile_config=tile_config, persistent_per_cu=choice["forward"], use_graph=True))
            if best is None or total < best[0]:
                best = (total, tile_config, choice)
        tile_config, per_cu = best[1], best[2]
        t_fwd = api.tuning(clear=True, tile_config=tile_config, persistent_per_cu=per_cu["forward"], use_graph=True)
        t_bwd = api.tuning(clear=True, tile_config=tile_config, persistent_per_cu=per_cu["backward"], use_graph=True)
        for _ in range(2):
            r.forward(gv.model, gv.views, tex, mul, t_fwd)
            r.backward(gt, gm, t_bwd)
        fwd = timed(lambda: r.forward(gv.model, gv.views, tex, mul, t_fwd), steps)
        bwd = timed(lambda: r.backward(gt, gm, t_bwd), steps)
        pixels = sum(v.width * v.height for v in views)
        progress(f"differentiable {key}: forward {fwd * 1e3:.3f} ms, backward {bwd * 1e3:.3f} ms")
        out[key] = {"views": len(views), "triangles": int(views[0].triangle_count), "texture": tex_size,
                    "forward_ms": round(fwd * 1e3, 3), "backward_ms": round(bwd * 1e3, 3),
                    "step_ms": round((fwd + bwd) * 1e3, 3),
                    "step_mpix_per_s": round(pixels / (fwd + bwd) / 1e6, 1),
                    "tile": ["16x16", "32x16", "32x32"][tile_config], "persistent_per_cu": per_cu}
    return out


def depth_only_form(calls: int = 20) -> dict:
    """model_renderDepth of the 1080p terrain into a device depth canvas, per synchronous call."""
    import torch
    from dfpsr_amd import api, scenes
    sc = scenes.terrain()
    sc.has_color = False
    g = api.GpuScene(sc)
    depth = torch.zeros((sc.height, sc.width), dtype=torch.float32, device="cuda")
    im = api.device_image(depth)
    for _ in range(3):
        depth.zero_()
        api.model_renderDepth(g.model, sc.instances[0], im, g.camera)
    total = 0.0
    for _ in range(calls):
        depth.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        api.model_renderDepth(g.model, sc.instances[0], im, g.camera)
        total += time.perf_counter() - t0
    ms = total / calls * 1e3
    return {"workload": "terrain 1920x1080, 130,050 triangles, depth canvas only", "ms_per_call": round(ms, 4),
            "mpix_per_s": round(sc.width * sc.height / (ms * 1e-3) / 1e6, 1),
            "mtri_per_s": round(int((sc.indices[:, 3] >= 0).sum() + sc.indices.shape[0]) / (ms * 1e-3) / 1e6, 1)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-baseline-only", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--no-depth", action="store_true")
    ap.add_argument("--tile-config", type=int, default=-1, help="renderer tile shape 0..2 (default: fastest in a sweep)")
    ap.add_argument("--graph", type=int, default=-1, help="1/0: replay the frame as a hipGraph (default: fastest)")
    ap.add_argument("--persistent", type=int, default=0, help="persistent raster workgroups per CU (0 = one per tile)")
    ap.add_argument("--raster-path", type=int, default=0, help="0 visibility raster where allowed, 1 in-order")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the untimed tuning sweep (with --tile-config and --graph, for profiling one tuning)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES between two device buffers three times (calibrates HBM counters)")
    args = ap.parse_args()
    if args.cpu_baseline_only:
        print(json.dumps(cpu_baseline()))
        return 0
    if args.copy_probe:
        import torch
        src = torch.ones(args.copy_probe, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "launches": 3}))
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
    color = torch.zeros((HEIGHT, WIDTH), dtype=torch.int32, device="cuda")
    depth = torch.zeros((HEIGHT, WIDTH), dtype=torch.float32, device="cuda")
    color_im, depth_im = api.device_image(color), api.device_image(depth)

    def make_renderer(tile_config, persistent, graph, path):
        r = api.renderer_create()
        # clear = image_fill of both canvases (colour 0, depth 0 = infinitely far) inside the frame.
        api.renderer_setTuning(r, api.tuning(tile_config=tile_config, persistent_per_cu=persistent, use_graph=graph,
                                             clear=True, raster_path=path))
        return r

    def step(r):
        api.renderer_begin(r, color_im, depth_im)
        gpu_scene.give(r)
        api.renderer_end(r)

    # Untimed sweep of the renderer's tunings; the timed run uses the fastest (or the flags').
    sweep = {}
    for tile_config in ((0, 1, 2) if not args.no_sweep else (max(args.tile_config, 0),)):
        for graph in ((False, True) if not args.no_sweep else (bool(max(args.graph, 0)),)):
            r = make_renderer(tile_config, args.persistent, graph, args.raster_path)
            for _ in range(3):
                step(r)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(10):
                step(r)
            torch.cuda.synchronize()
            sweep[(tile_config, graph)] = (time.perf_counter() - t0) / 10 * 1e3
    best = min(sweep, key=sweep.get)
    tile_config = best[0] if args.tile_config < 0 else args.tile_config
    graph = best[1] if args.graph < 0 else bool(args.graph)
    progress("plain sweep (ms/frame): " + ", ".join(f"tile{k[0]}{'+graph' if k[1] else ''}={v:.3f}" for k, v in sweep.items()))
    renderer = make_renderer(tile_config, args.persistent, graph, args.raster_path)

    for _ in range(args.warmup):
        step(renderer)
    # Timed region: the K frames as an application draws them.
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(renderer)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    stats = renderer.stats()
    progress(f"plain: {elapsed / args.steps * 1e3:.3f} ms per frame (untimed by events)")
    # The same K frames again with the renderer's HIP events on its own stream around each stage;
    # reading them back after every frame costs host time, so this pass is not the one timed above.
    renderer.set_timing(True)
    raster_ms = []
    stage_ms = np.zeros(5)
    for _ in range(args.steps):
        step(renderer)
        t = renderer.timing()
        raster_ms.append(float(t[3]))
        stage_ms += t
    renderer.set_timing(False)
    ms_per_step = elapsed / args.steps * 1e3
    raster_avg_ms = float(np.mean(raster_ms))
    # The algorithmic bytes of the whole frame over the dominant kernel's average launch: the
    # frame's bytes as the roofline's unit of work, the tile raster as the kernel that bounds it.
    achieved = FRAME_BYTES / (raster_avg_ms * 1e-3) / 1e9
    device_frame_ms = float(stage_ms[4]) / args.steps
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
                   "width": WIDTH, "height": HEIGHT, "triangles": TRIANGLES, "parallelism": "single GPU",
                   "tuning": {"tile": ["16x16", "32x16", "32x32"][tile_config], "graph": bool(graph),
                              "persistent_per_cu": args.persistent,
                              "raster": "in-order" if args.raster_path else "visibility"}},
        "plain_sweep_ms": {f"{['16x16', '32x16', '32x32'][k[0]]}{'+graph' if k[1] else ''}": round(v, 4)
                           for k, v in sweep.items()},
        "mtri_per_s": round(TRIANGLES * args.steps / elapsed / 1e6, 3),
        "frames_per_s": round(args.steps / elapsed, 2),
        "frame_gbs": round(FRAME_BYTES / (ms_per_step * 1e-3) / 1e9, 2),
        "stage_ms": {k: round(float(v) / args.steps, 4) for k, v in
                     zip(["project", "setup", "binning", "raster", "device_frame"], stage_ms)},
        "draw_commands": int(stats[0]),
        "tile_pairs": int(stats[1]),
        "roofline": {"bound": "hbm", "kernel": "k_raster_deferred" if not args.raster_path else "k_raster", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(),
                     "algorithmic_bytes": FRAME_BYTES, "avg_launch_ms": round(raster_avg_ms, 5),
                     # The same frame bytes over the device time of the whole frame (every kernel
                     # and the gaps between them), and the raster kernel's own share of the bytes
                     # (canvas pixels and sampled texels) over its launch.
                     "frame": {"bytes": FRAME_BYTES, "device_ms": round(device_frame_ms, 5),
                               "achieved": round(FRAME_BYTES / (device_frame_ms * 1e-3) / 1e9, 2),
                               "frac": round(FRAME_BYTES / (device_frame_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
                     "raster_share": {"bytes": RASTER_BYTES, "achieved": round(RASTER_BYTES / (raster_avg_ms * 1e-3) / 1e9, 2),
                                      "frac": round(RASTER_BYTES / (raster_avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}},
        "cpu_baseline": None,
    }
    del gpu_scene, renderer, color, depth, color_im, depth_im
    if not args.no_batch:
        line["batched"] = batched_form()
        torch.cuda.empty_cache()
    if not args.no_diff:
        line["differentiable"] = differentiable_form()
        torch.cuda.empty_cache()
    if not args.no_depth:
        line["depth_only"] = depth_only_form()
        torch.cuda.empty_cache()
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
