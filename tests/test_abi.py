"""CPU-side checks of the C ABI: the library loads, exports every function include/dfpsr_hip.h
declares, and validates arguments that never reach the device."""
import ctypes as C

import numpy as np
import pytest

from dfpsr_amd import _lib


@pytest.fixture(scope="module")
def lib():
    return _lib.lib()


def test_exports_every_declared_symbol(lib):
    names = _lib.declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version(lib):
    assert b"gfx950" in lib.dfpsr_hip_version()


def test_model_argument_checks(lib):
    m = C.c_void_p()
    assert lib.dfpsr_hip_model_create(C.byref(m), 7) == _lib.ERROR_ARGUMENT
    assert b"filter" in lib.dfpsr_hip_last_error()
    assert lib.dfpsr_hip_model_create(C.byref(m), 0) == 0
    pts = np.zeros((3, 3), np.float32)
    assert lib.dfpsr_hip_model_set_points(m, 3, pts.ctypes.data_as(C.c_void_p)) == 0
    idx = np.array([[0, 1, 5, -1]], np.int32)
    col = np.ones((1, 16), np.float32)
    tex = np.zeros((1, 16), np.float32)
    rc = lib.dfpsr_hip_model_add_part(m, 1, idx.ctypes.data_as(C.c_void_p), col.ctypes.data_as(C.c_void_p),
                                      tex.ctypes.data_as(C.c_void_p), None, None)
    assert rc == _lib.ERROR_ARGUMENT and b"out of range" in lib.dfpsr_hip_last_error()
    idx[0, 2] = 2
    rc = lib.dfpsr_hip_model_add_part(m, 1, idx.ctypes.data_as(C.c_void_p), col.ctypes.data_as(C.c_void_p),
                                      tex.ctypes.data_as(C.c_void_p), None, None)
    assert rc == 0
    lib.dfpsr_hip_model_destroy(m)


def test_model_bound_includes_origin(lib):
    """ModelImpl's bound starts at the origin (Model.cpp:281-311)."""
    m = C.c_void_p()
    lib.dfpsr_hip_model_create(C.byref(m), 0)
    pts = np.array([[1, 2, 3], [4, 5, 6]], np.float32)
    lib.dfpsr_hip_model_set_points(m, 2, pts.ctypes.data_as(C.c_void_p))
    mn = np.zeros(3, np.float32)
    mx = np.zeros(3, np.float32)
    assert lib.dfpsr_hip_model_bound(m, mn.ctypes.data_as(C.c_void_p), mx.ctypes.data_as(C.c_void_p)) == 0
    assert mn.tolist() == [0, 0, 0] and mx.tolist() == [4, 5, 6]
    lib.dfpsr_hip_model_destroy(m)


def test_null_handles(lib):
    assert lib.dfpsr_hip_renderer_end(None) == _lib.ERROR_ARGUMENT
    assert lib.dfpsr_hip_renderer_begin(None, None, None) == _lib.ERROR_ARGUMENT
    assert lib.dfpsr_hip_texture_generate_pyramid(None) == _lib.ERROR_ARGUMENT
    assert lib.dfpsr_hip_texture_create(None, 4, 4, 1) == _lib.ERROR_ARGUMENT
    t = C.c_void_p()
    assert lib.dfpsr_hip_texture_create(C.byref(t), 0, 4, 1) == _lib.ERROR_ARGUMENT
    assert lib.dfpsr_hip_texture_create(C.byref(t), 4, 4, 0) == _lib.ERROR_ARGUMENT


def test_points_replaced_by_fewer_are_refused_before_drawing(lib):
    """model_set_points after model_add_part can leave polygons naming missing points; a draw
    must refuse that before it reaches the device (the depth-only draw is checked here; the
    renderer, batch and differentiable paths in tests/test_gpu_parity.py)."""
    m = C.c_void_p()
    assert lib.dfpsr_hip_model_create(C.byref(m), 0) == 0
    pts = np.array([[0, 0, 1], [1, 0, 1], [0, 1, 1]], np.float32)
    assert lib.dfpsr_hip_model_set_points(m, 3, pts.ctypes.data_as(C.c_void_p)) == 0
    idx = np.array([[0, 1, 2, -1]], np.int32)
    col = np.ones((1, 16), np.float32)
    tex = np.zeros((1, 16), np.float32)
    assert lib.dfpsr_hip_model_add_part(m, 1, idx.ctypes.data_as(C.c_void_p), col.ctypes.data_as(C.c_void_p),
                                        tex.ctypes.data_as(C.c_void_p), None, None) == 0
    assert lib.dfpsr_hip_model_set_points(m, 2, pts.ctypes.data_as(C.c_void_p)) == 0
    canvas = np.zeros((8, 8), np.float32)
    img = _lib.Image(canvas.ctypes.data, 8, 8, 32, 0)
    cam = _lib.Camera(1, (C.c_float * 12)(0, 0, -2, 1, 0, 0, 0, 1, 0, 0, 0, 1), 8.0, 8.0, 1.0, 0.01, 100.0)
    t = (C.c_float * 12)(0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1)
    rc = lib.dfpsr_hip_model_render_depth(m, t, C.byref(img), C.byref(cam))
    assert rc == _lib.ERROR_ARGUMENT and b"does not have" in lib.dfpsr_hip_last_error()
    assert not canvas.any()
    lib.dfpsr_hip_model_destroy(m)


def test_scene_give_arguments_prepared_once_and_remade_on_replacement():
    """GpuScene.give's renderer_giveTask arguments: the instances' float32 transforms and the
    camera reference are made once, reused while the scene's instance array and the camera object
    stay the same, and made again when either is replaced (host logic only, no device work)."""
    from types import SimpleNamespace
    from dfpsr_amd import api
    g = api.GpuScene.__new__(api.GpuScene)
    g.scene = SimpleNamespace(instances=np.stack([np.eye(4, dtype=np.float64)[:3].reshape(-1)] * 2))
    axes = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)
    g.camera = api.camera_create(True, (0.0, 0.0, -5.0) + axes, 64, 48, 1.0)
    g._prepared = None
    first = g._prepared_arguments()
    transforms, pointers = first[1], first[2]
    assert len(transforms) == 2 and all(t.dtype == np.float32 and t.flags.c_contiguous for t in transforms)
    assert [p.value for p in pointers] == [t.ctypes.data for t in transforms]
    np.testing.assert_array_equal(transforms[1], g.scene.instances[1].astype(np.float32))
    assert g._prepared_arguments() is first
    g.camera = api.camera_create(True, (0.0, 1.0, -5.0) + axes, 64, 48, 1.0)
    second = g._prepared_arguments()
    assert second is not first
    g.scene.instances = g.scene.instances[:1] * 2.0
    third = g._prepared_arguments()
    assert third is not second and len(third[1]) == 1
    np.testing.assert_array_equal(third[1][0], (g.scene.instances[0]).astype(np.float32))
