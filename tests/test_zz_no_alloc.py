"""The scratch-free depth-only draw, dfpsr_hip_model_render_depth_noalloc, against the
allocating dfpsr_hip_model_render_depth on the same inputs: bit-identical depth canvases on
two calls in a row, the second with a smaller model and canvas, and its refusals (host
canvas, a model with changes not yet on the device, more points than its scratch array)."""
import ctypes as C

import numpy as np
import pytest

try:  # import torch at collection time: the first import on a fresh machine is slow
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

from dfpsr_amd import _lib
from dfpsr_amd import scenes as S

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 18


@pytest.fixture(scope="module")
def api():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from dfpsr_amd import api as A
    l = A.lib()
    l.dfpsr_hip_model_render_depth_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    l.dfpsr_hip_model_render_depth_noalloc.restype = C.c_int
    return A


def draw(api, g, scene, noalloc):
    import torch
    depth = np.zeros((scene.height, scene.width), np.float32) if scene.initial_depth is None else scene.initial_depth.astype(np.float32)
    dev = torch.from_numpy(np.ascontiguousarray(depth)).cuda()
    im = api.device_image(dev)
    for t in scene.instances:
        tr = np.ascontiguousarray(t, dtype=np.float32)
        if noalloc:
            rc = api.lib().dfpsr_hip_model_render_depth_noalloc(g.model.handle, tr.ctypes.data_as(C.c_void_p), C.byref(im),
                                                                C.byref(g.camera))
            api.check(rc)
        else:
            api.model_renderDepth(g.model, tr, im, g.camera)
    torch.cuda.synchronize()
    return dev.cpu().numpy()


def test_noalloc_matches_allocating_entry_point_on_two_calls(api):
    big = S.terrain()
    big.has_color = False
    small = S.cube_soup(width=200, height=150, triangles=700, seed=11, spread=0.4)
    small.has_color = False
    gb, gs = api.GpuScene(big), api.GpuScene(small)
    # Make both models resident with the allocating call first (the scratch-free call uploads nothing).
    want_big = draw(api, gb, big, False)
    want_small = draw(api, gs, small, False)
    assert (want_big > 0).mean() > 0.5 and (want_small > 0).any()
    print("[gpu] noalloc: 1080p terrain, then a 200x150 soup", flush=True)
    got_big = draw(api, gb, big, True)
    got_small = draw(api, gs, small, True)
    assert np.array_equal(got_big.view(np.uint32), want_big.view(np.uint32))
    assert np.array_equal(got_small.view(np.uint32), want_small.view(np.uint32))
    # Once more in the other order: the shared array keeps the previous call's points.
    assert np.array_equal(draw(api, gs, small, True).view(np.uint32), want_small.view(np.uint32))


def test_noalloc_refusals(api):
    sc = S.cube_soup(width=32, height=24, triangles=20, seed=2)
    g = api.GpuScene(sc)
    t = np.ascontiguousarray(sc.instances[0], dtype=np.float32)
    fn = api.lib().dfpsr_hip_model_render_depth_noalloc
    host = np.zeros((sc.height, sc.width), np.float32)
    him = api.host_image(host)
    assert fn(g.model.handle, t.ctypes.data_as(C.c_void_p), C.byref(him), C.byref(g.camera)) == _lib.ERROR_ARGUMENT
    assert b"device memory" in api.lib().dfpsr_hip_last_error()
    import torch
    dev = torch.zeros((sc.height, sc.width), dtype=torch.float32, device="cuda")
    dim = api.device_image(dev)
    # Never drawn: its arrays are not on the device yet.
    assert fn(g.model.handle, t.ctypes.data_as(C.c_void_p), C.byref(dim), C.byref(g.camera)) == _lib.ERROR_STATE
    assert not dev.any().item()
    big = api.model_create()
    pts = np.zeros((MAX_POINTS + 1, 3), np.float32)
    pts[:, 2] = 1.0
    api.model_setPoints(big, pts)
    api.model_addPart(big, np.array([[0, 1, 2, -1]], np.int32), np.ones((1, 16), np.float32), np.zeros((1, 16), np.float32))
    assert fn(big.handle, t.ctypes.data_as(C.c_void_p), C.byref(dim), C.byref(g.camera)) == _lib.ERROR_ARGUMENT
    assert b"scratch" in api.lib().dfpsr_hip_last_error()
