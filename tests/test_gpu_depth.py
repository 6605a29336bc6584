"""Depth-only drawing (model_renderDepth) on the gfx950 library, through its C ABI, against
the CPU oracle and the reference's depth canvases (tests/golden/depth_scenes.npz).

The bar is bit-exact depth: rows keep the reference's per-pixel float adds and the
keep-nearest write is order independent.
"""
import os

import numpy as np
import pytest

try:  # import torch at collection time: the first import on a fresh machine is slow
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

from dfpsr_amd import scenes as S
from oracle import harness as H
import fixture_scenes

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def api():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from dfpsr_amd import api as A
    A.lib()
    return A


def assert_same_depth(name, got, want):
    bad = got.view(np.uint32) != want.view(np.uint32)
    assert not bad.any(), f"{name}: {int(bad.sum())} depth pixels differ, first at {np.argwhere(bad)[:3].tolist()}"


@pytest.mark.parametrize("scene", fixture_scenes.depth_scenes(), ids=lambda s: s.name)
def test_depth_only_matches_reference(api, scene):
    print(f"depth-only {scene.name} {scene.width}x{scene.height}", flush=True)
    want = np.load(os.path.join(GOLDEN, "depth_scenes.npz"))[scene.name]
    assert_same_depth(scene.name, api.render_scene_depth(scene), want)
    assert_same_depth(scene.name, H.oracle_render_depth(scene), want)


def test_depth_only_device_canvas(api):
    sc = fixture_scenes.depth_scenes()[6]
    print(f"depth-only on a device canvas {sc.name}", flush=True)
    assert_same_depth(sc.name, api.render_scene_depth(sc, on_device=True), H.oracle_render_depth(sc))


@pytest.mark.parametrize("height", [47, 61])
def test_depth_only_odd_height(api, height):
    """The last row bound rounds up to an even row past the canvas; that row is not drawn."""
    sc = S.cube_soup(width=70, height=height, triangles=400, seed=height, spread=0.6, cam_distance=1.2, name=f"odd{height}")
    sc.has_color = False
    print(f"depth-only odd height {sc.name}", flush=True)
    got = api.render_scene_depth(sc)
    assert (got > 0).sum() > 100
    assert_same_depth(sc.name, got, H.oracle_render_depth(sc))


def test_depth_only_terrain_1080p(api):
    """The 1920x1080 terrain (130 050 triangles) with its depth alone."""
    sc = S.terrain()
    sc.has_color = False
    print("depth-only terrain 1920x1080", flush=True)
    got = api.render_scene_depth(sc)
    assert (got > 0).mean() > 0.5
    assert_same_depth(sc.name, got, H.oracle_render_depth(sc))


def test_depth_only_without_canvas_draws_nothing(api):
    sc = fixture_scenes.depth_scenes()[0]
    g = api.GpuScene(sc)
    api.model_renderDepth(g.model, sc.instances[0], api.host_image(None), g.camera)


@pytest.mark.parametrize("index", [0, 1])
def test_depth_only_special_canvas_values(api, index):
    """A canvas preloaded with -0, +0, negative numbers, NaN of both signs, +inf and tiny and
    huge positives: the keep-nearest write follows the reference's strict float comparison
    on each of them (perspective keeps the greater value, orthogonal the smaller)."""
    sc = fixture_scenes.special_depth_canvas(fixture_scenes.depth_scenes()[index], 100 + index)
    print(f"depth-only special canvas {sc.name} perspective={sc.perspective}", flush=True)
    got = api.render_scene_depth(sc)
    want = H.oracle_render_depth(sc)
    assert (got.view(np.uint32) != sc.initial_depth.view(np.uint32)).sum() > 50
    assert_same_depth(sc.name, got, want)
