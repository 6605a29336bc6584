"""Autotuner of the batched form: picks the raster tile shape and the number of persistent
workgroups per CU for this device and stores the choice as JSON.

    python -m dfpsr_amd.autotune [--jobs 4096] [--repeats 3] [--out dfpsr_amd/tuning_gfx950.json]

The sweep itself runs in the library (dfpsr_hip_batch_autotune): every tile shape
(16x16, 32x16, 32x32) x persistent workgroups per CU in {0, 1, 2, 4, 8}, graph replay on,
timed as the median of whole batch renders.  The workload is the uneven-job batch of
dfpsr_amd.scenes.batch_jobs (canvases 64..640 x 48..480, 100..20,000 triangles per job,
cube soups and terrain patches, shared 256x256 textures on half the jobs).
"""
from __future__ import annotations

import argparse
import json
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PATH = os.path.join(HERE, "tuning_gfx950.json")
PERSISTENT_STEPS = (0, 1, 2, 4, 8)
TILE_SHAPES = ("16x16", "32x16", "32x32")


def run(job_count: int = 4096, repeats: int = 3, seed: int = 2) -> dict:
    import torch
    from . import api, scenes
    t0 = time.time()
    jobs = scenes.batch_jobs(count=job_count, seed=seed)
    gpu = api.GpuJobs(jobs)
    batch = api.Batch()
    best, ms = batch.autotune(gpu.jobs, repeats, api.tuning(clear=True))
    torch.cuda.synchronize()
    props = torch.cuda.get_device_properties(0)
    return {
        "device": props.name,
        "arch": getattr(props, "gcnArchName", ""),
        "compute_units": props.multi_processor_count,
        "workload": {"jobs": job_count, "seed": seed, "triangles": int(sum(j.triangle_count for j in jobs)),
                     "pixels": int(sum(j.width * j.height for j in jobs))},
        "repeats": repeats,
        "best": {"tile_config": int(best.tile_config), "tile": TILE_SHAPES[best.tile_config],
                 "persistent_per_cu": int(best.persistent_per_cu), "use_graph": 1},
        "ms": {f"{TILE_SHAPES[tc]}/p{PERSISTENT_STEPS[p]}": round(float(ms[tc, p]), 4)
               for tc in range(3) for p in range(len(PERSISTENT_STEPS))},
        "seconds": round(time.time() - t0, 1),
    }


def load(path: str = DEFAULT_PATH):
    """The stored choice as (tile_config, persistent_per_cu), or None when there is none."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        best = json.load(f)["best"]
    return int(best["tile_config"]), int(best["persistent_per_cu"])


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--jobs", type=int, default=4096)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--out", default=DEFAULT_PATH)
    args = ap.parse_args()
    result = run(args.jobs, args.repeats, args.seed)
    with open(args.out, "w") as f:
        json.dump(result, f, indent=1)
    print(json.dumps(result["best"]), flush=True)


if __name__ == "__main__":
    main()
