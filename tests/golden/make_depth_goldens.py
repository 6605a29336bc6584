"""Writes tests/golden/depth_scenes.npz: the depth canvas that the reference's model_renderDepth
(modelAPI.cpp:202-206) leaves for every scene of fixture_scenes.depth_scenes().

Run in the build container only (it needs oracle/_ref/libdfpsr_ref.so, built from the reference
sources by `make -f oracle/Makefile.ref`):

    python tests/golden/make_depth_goldens.py
"""
from __future__ import annotations

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.dirname(HERE)]

from oracle import harness  # noqa: E402

harness.load_ref()  # load before any large numpy work (its static heap set-up needs it)

import numpy as np  # noqa: E402
import fixture_scenes  # noqa: E402


def main():
    out = {}
    for sc in fixture_scenes.depth_scenes():
        out[sc.name] = harness.ref_render_depth(sc)
    np.savez_compressed(os.path.join(HERE, "depth_scenes.npz"), **out)
    print(f"wrote {len(out)} depth canvases")


if __name__ == "__main__":
    main()
