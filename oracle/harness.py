"""Test infrastructure only: ctypes bridges to the CPU oracle (oracle/_build/liboracle.so)
and, in the build container, to the reference renderer compiled from its own sources
(oracle/_ref/libdfpsr_ref.so).  Both take the same scene_desc (oracle/scene_desc.h).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this module,
and only as the checker or the timed CPU baseline -- never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# DFPSR_ORACLE_LIBRARY names another build of the oracle (the sanitized one, for instance).
ORACLE_SO = os.environ.get("DFPSR_ORACLE_LIBRARY") or os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libdfpsr_ref.so")


# ProjectedPoint (ProjectedPoint.h:33-49) as 40-byte records: cs xyz, is xy, flat xy.
PROJECTED_DTYPE = np.dtype({"names": ["cs", "is", "flat"], "formats": [(np.float32, 3), (np.float32, 2), (np.int64, 2)],
                            "offsets": [0, 12, 24], "itemsize": 40})


# Occluder calls of a scene: (kind, instance it comes before, box minimum, maximum, transform).
OCCLUDER_KINDS = {"box": 1, "top_rows": 2, "existing": 3, "query": 4}


class SceneDesc(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32),
        ("has_color", C.c_int32), ("has_depth", C.c_int32),
        ("initial_color", C.c_void_p), ("initial_depth", C.c_void_p),
        ("cam_location", C.c_float * 12),
        ("perspective", C.c_int32),
        ("width_slope", C.c_float), ("near_clip", C.c_float), ("far_clip", C.c_float),
        ("filter", C.c_int32),
        ("point_count", C.c_int32), ("points", C.c_void_p),
        ("part_count", C.c_int32),
        ("part_polygon_start", C.c_void_p), ("part_diffuse", C.c_void_p), ("part_light", C.c_void_p),
        ("indices", C.c_void_p), ("colors", C.c_void_p), ("texcoords", C.c_void_p),
        ("texture_count", C.c_int32), ("texture_dims", C.c_void_p), ("texture_level0", C.c_void_p),
        ("instance_count", C.c_int32), ("instance_transforms", C.c_void_p),
        ("threaded", C.c_int32),
        ("projected", C.c_void_p),
        ("occluder_count", C.c_int32), ("occluders", C.c_void_p), ("visible_out", C.c_void_p),
    ]


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def make_desc(scene, threaded: bool = True) -> Tuple[SceneDesc, list]:
    """Build a scene_desc for `scene`; the returned list keeps the arrays alive."""
    keep = []

    def arr(x, dt):
        a = np.ascontiguousarray(x, dtype=dt)
        keep.append(a)
        return a

    d = SceneDesc()
    d.width, d.height = scene.width, scene.height
    d.has_color, d.has_depth = int(scene.has_color), int(scene.has_depth)
    d.initial_color = _ptr(arr(scene.initial_color, np.uint32)) if scene.initial_color is not None else None
    d.initial_depth = _ptr(arr(scene.initial_depth, np.float32)) if scene.initial_depth is not None else None
    d.cam_location[:] = [float(v) for v in np.asarray(scene.cam_location, dtype=np.float32)]
    d.perspective = int(scene.perspective)
    d.width_slope, d.near_clip, d.far_clip = scene.width_slope, scene.near_clip, scene.far_clip
    d.filter = scene.filter
    d.point_count = scene.points.shape[0]
    d.points = _ptr(arr(scene.points, np.float32))
    d.part_count = len(scene.parts)
    starts = [p.polygon_start for p in scene.parts] + [scene.parts[-1].polygon_end if scene.parts else 0]
    d.part_polygon_start = _ptr(arr(starts, np.int32))
    d.part_diffuse = _ptr(arr([p.diffuse for p in scene.parts] or [0], np.int32))
    d.part_light = _ptr(arr([p.light for p in scene.parts] or [0], np.int32))
    d.indices = _ptr(arr(scene.indices, np.int32))
    d.colors = _ptr(arr(scene.colors, np.float32))
    d.texcoords = _ptr(arr(scene.texcoords, np.float32))
    d.texture_count = len(scene.textures)
    dims = []
    ptrs = (C.c_void_p * max(1, len(scene.textures)))()
    for i, t in enumerate(scene.textures):
        dims += [t.width, t.height, t.resolutions]
        lv = arr(t.level0, np.uint32)
        ptrs[i] = lv.ctypes.data
    keep.append(ptrs)
    d.texture_dims = _ptr(arr(dims or [0], np.int32))
    d.texture_level0 = C.cast(ptrs, C.c_void_p)
    d.instance_count = scene.instances.shape[0]
    d.instance_transforms = _ptr(arr(scene.instances, np.float32))
    d.threaded = int(threaded)
    proj = getattr(scene, "projected", None)
    if proj is not None:
        a = np.ascontiguousarray(proj)
        assert a.dtype == PROJECTED_DTYPE and a.shape == (scene.indices.shape[0], 3)
        keep.append(a)
        d.projected = a.ctypes.data_as(C.c_void_p)
    ops = getattr(scene, "occluders", None)
    if ops:
        table = np.zeros((len(ops), 20), np.float32)
        for k, op in enumerate(ops):
            table[k, 0] = OCCLUDER_KINDS[op[0]]
            table[k, 1] = op[1]
            if op[0] in ("box", "query"):
                table[k, 2:5], table[k, 5:8], table[k, 8:20] = op[2], op[3], op[4]
        keep.append(table)
        d.occluder_count = len(ops)
        d.occluders = table.ctypes.data_as(C.c_void_p)
        visible = np.full(max(1, sum(op[0] == "query" for op in ops)), -1, np.int32)
        keep.append(visible)
        d.visible_out = visible.ctypes.data_as(C.c_void_p)
    return d, keep


_libs = {}


def _load(path: str, name: str):
    if name not in _libs:
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built")
        lib = C.CDLL(path)
        _libs[name] = lib
    return _libs[name]


def ref_available() -> bool:
    return os.path.exists(REF_SO)


_ref_exit_registered = False


def load_ref():
    global _ref_exit_registered
    lib = _load(REF_SO, "ref")
    if not _ref_exit_registered and hasattr(lib, "ref_terminate"):
        import atexit
        atexit.register(lib.ref_terminate)
        _ref_exit_registered = True
    lib.ref_render_scene.argtypes = [C.POINTER(SceneDesc), C.c_void_p, C.c_void_p]
    lib.ref_time_scene.argtypes = [C.POINTER(SceneDesc), C.c_int32, C.c_int32, C.POINTER(C.c_double)]
    lib.ref_render_depth_scene.argtypes = [C.POINTER(SceneDesc), C.c_void_p]
    return lib


def load_oracle():
    lib = _load(ORACLE_SO, "oracle")
    lib.oracle_render_scene.argtypes = [C.POINTER(SceneDesc), C.c_int32, C.c_void_p, C.c_void_p]
    lib.oracle_render_scene.restype = C.c_int32
    lib.oracle_render_depth_scene.argtypes = [C.POINTER(SceneDesc), C.c_void_p]
    lib.oracle_render_depth_scene.restype = C.c_int32
    lib.oracle_time_scene.argtypes = [C.POINTER(SceneDesc), C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double)]
    lib.oracle_time_scene.restype = C.c_int32
    lib.oracle_texture_query.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_int32,
                                         C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_float]
    lib.oracle_texture_query.restype = C.c_uint32
    lib.oracle_mix_colors.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    lib.oracle_mix_colors.restype = C.c_uint32
    lib.oracle_pack_saturated.argtypes = [C.c_float] * 4
    lib.oracle_pack_saturated.restype = C.c_uint32
    lib.oracle_reciprocal.argtypes = [C.c_int32, C.c_void_p, C.c_void_p]
    lib.oracle_rasterize.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
    lib.oracle_texture_pyramid.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32,
                                           C.c_void_p, C.c_void_p, C.c_void_p]
    lib.oracle_texture_pyramid.restype = C.c_int32
    lib.oracle_sample_quads.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p]
    lib.oracle_project.argtypes = [C.POINTER(SceneDesc), C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.oracle_occlusion_grid.argtypes = [C.POINTER(SceneDesc), C.c_void_p, C.c_void_p, C.c_void_p]
    lib.oracle_occlusion_grid.restype = C.c_int32
    lib.oracle_commands.argtypes = [C.POINTER(SceneDesc), C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.oracle_commands.restype = C.c_int32
    return lib


def oracle_diff_forward(scene, texture: np.ndarray, multipliers=None):
    """Differentiable float shader on the oracle's raster: (colour HxWx4, visibility HxW,
    depth HxW, weights HxWx5)."""
    lib = load_oracle()
    lib.oracle_diff_forward.argtypes = [C.POINTER(SceneDesc), C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.oracle_diff_forward.restype = C.c_int32
    d, keep = make_desc(scene, False)
    tex = np.ascontiguousarray(texture, dtype=np.float32)
    mul = None if multipliers is None else np.ascontiguousarray(multipliers, dtype=np.float32)
    h, w = scene.height, scene.width
    color = np.zeros((h, w, 4), np.float32)
    vis = np.zeros((h, w), np.int32)
    depth = np.zeros((h, w), np.float32)
    weights = np.zeros((h, w, 5), np.float32)
    rc = lib.oracle_diff_forward(C.byref(d), ptr(tex), int(tex.shape[1]).bit_length() - 1, int(tex.shape[0]).bit_length() - 1,
                                 None if mul is None else ptr(mul), ptr(color), ptr(vis), ptr(depth), ptr(weights))
    if rc < 0:
        raise RuntimeError("oracle_diff_forward failed")
    return color, vis, depth, weights


def oracle_diff_backward(scene, texture, multipliers, visibility, weights, grad_color, point_count):
    """Adjoint of oracle_diff_forward in float64: (dL/dtexture, dL/dmultipliers)."""
    lib = load_oracle()
    lib.oracle_diff_backward.argtypes = [C.POINTER(SceneDesc), C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    d, keep = make_desc(scene, False)
    tex = np.ascontiguousarray(texture, dtype=np.float32)
    mul = None if multipliers is None else np.ascontiguousarray(multipliers, dtype=np.float32)
    gt = np.zeros(tex.shape, np.float64)
    gm = np.zeros((point_count, 3), np.float64)
    lib.oracle_diff_backward(C.byref(d), ptr(tex), int(tex.shape[1]).bit_length() - 1, int(tex.shape[0]).bit_length() - 1,
                             None if mul is None else ptr(mul), ptr(np.ascontiguousarray(visibility, np.int32)),
                             ptr(np.ascontiguousarray(weights, np.float32)), ptr(np.ascontiguousarray(grad_color, np.float32)),
                             ptr(gt), ptr(gm))
    return gt, gm


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def oracle_rasterize(corners, left, top, width, height):
    lib = load_oracle()
    c = np.ascontiguousarray(corners, dtype=np.int64)
    rows = np.zeros((height, 2), dtype=np.int32)
    lib.oracle_rasterize(ptr(c), left, top, width, height, ptr(rows))
    return rows


def oracle_commands(scene):
    """Draw commands of a frame: (source triangle, flat corners (n,3,2), pixel bound (n,4))."""
    lib = load_oracle()
    d, keep = make_desc(scene, False)
    n = lib.oracle_commands(C.byref(d), 0, None, None, None)
    tri = np.zeros(n, dtype=np.int32)
    flat = np.zeros((n, 3, 2), dtype=np.int64)
    bound = np.zeros((n, 4), dtype=np.int32)
    lib.oracle_commands(C.byref(d), n, ptr(tri), ptr(flat), ptr(bound))
    return tri, flat, bound


def oracle_project(scene, transform, points):
    lib = load_oracle()
    d, keep = make_desc(scene, False)
    p = np.ascontiguousarray(points, dtype=np.float32)
    t = np.ascontiguousarray(transform, dtype=np.float32)
    n = p.shape[0]
    cs = np.zeros((n, 3), np.float32)
    is_ = np.zeros((n, 2), np.float32)
    flat = np.zeros((n, 2), np.int64)
    lib.oracle_project(C.byref(d), ptr(t), n, ptr(p), ptr(cs), ptr(is_), ptr(flat))
    return cs, is_, flat


def _outputs(scene):
    color = np.zeros((scene.height, scene.width), dtype=np.uint32)
    depth = np.zeros((scene.height, scene.width), dtype=np.float32)
    return color, depth


def ref_render(scene, threaded: bool = True):
    lib = load_ref()
    d, keep = make_desc(scene, threaded)
    color, depth = _outputs(scene)
    rc = lib.ref_render_scene(C.byref(d), color.ctypes.data_as(C.c_void_p), depth.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise RuntimeError("reference renderer failed")
    return color, depth


def _visible_out(scene, keep):
    ops = getattr(scene, "occluders", None)
    if not ops:
        return np.zeros(0, np.int32)
    n = sum(op[0] == "query" for op in ops)
    return keep[-1][:n].copy()


def ref_render_queries(scene):
    """The compiled reference's frame with the scene's occluder calls: (colour, depth, the
    answers of its renderer_isBoxVisible calls)."""
    lib = load_ref()
    d, keep = make_desc(scene, True)
    color, depth = _outputs(scene)
    rc = lib.ref_render_scene(C.byref(d), color.ctypes.data_as(C.c_void_p), depth.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise RuntimeError("reference renderer failed")
    return color, depth, _visible_out(scene, keep)


def oracle_render_queries(scene, threads: int = 1):
    """The oracle's frame with the scene's occluder calls: (colour, depth, box answers)."""
    lib = load_oracle()
    d, keep = make_desc(scene, threads > 1)
    color, depth = _outputs(scene)
    rc = lib.oracle_render_scene(C.byref(d), threads, color.ctypes.data_as(C.c_void_p), depth.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise RuntimeError(f"oracle failed with {rc}")
    return color, depth, _visible_out(scene, keep)


def oracle_occlusion_grid(scene):
    """(grid of ceil(H/16) x ceil(W/16) cells, whether the frame had occluders)."""
    lib = load_oracle()
    d, keep = make_desc(scene, True)
    gw, gh = (scene.width + 15) // 16, (scene.height + 15) // 16
    grid = np.zeros((gh, gw), np.float32)
    w, h = C.c_int32(), C.c_int32()
    occluded = lib.oracle_occlusion_grid(C.byref(d), grid.ctypes.data_as(C.c_void_p), C.byref(w), C.byref(h))
    assert (w.value, h.value) == (gw, gh)
    return grid, bool(occluded)


def oracle_render(scene, threads: int = 1):
    """Render with the oracle; threads > 1 splits rows into 12 slices like CommandQueue::execute."""
    lib = load_oracle()
    d, keep = make_desc(scene, threads > 1)
    color, depth = _outputs(scene)
    rc = lib.oracle_render_scene(C.byref(d), threads, color.ctypes.data_as(C.c_void_p), depth.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise RuntimeError(f"oracle failed with {rc}")
    return color, depth


def _initial_depth(scene):
    if scene.initial_depth is None:
        return np.zeros((scene.height, scene.width), dtype=np.float32)
    return np.ascontiguousarray(scene.initial_depth, dtype=np.float32).copy()


def ref_render_depth(scene):
    """model_renderDepth of every instance with the compiled reference (depth canvas only)."""
    lib = load_ref()
    d, keep = make_desc(scene, False)
    d.has_depth = 1
    depth = _initial_depth(scene)
    if lib.ref_render_depth_scene(C.byref(d), depth.ctypes.data_as(C.c_void_p)) != 0:
        raise RuntimeError("reference depth renderer failed")
    return depth


def oracle_render_depth(scene):
    """model_renderDepth of every instance with the oracle (depth canvas only)."""
    lib = load_oracle()
    d, keep = make_desc(scene, False)
    d.has_depth = 1
    depth = _initial_depth(scene)
    rc = lib.oracle_render_depth_scene(C.byref(d), depth.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise RuntimeError(f"oracle failed with {rc}")
    return depth


def oracle_time(scene, threads: int, warmup: int, iterations: int) -> float:
    lib = load_oracle()
    d, keep = make_desc(scene, threads > 1)
    out = C.c_double()
    rc = lib.oracle_time_scene(C.byref(d), threads, warmup, iterations, C.byref(out))
    if rc != 0:
        raise RuntimeError(f"oracle timing failed with {rc}")
    return out.value


def ref_time(scene, warmup: int, iterations: int) -> float:
    lib = load_ref()
    d, keep = make_desc(scene, True)
    out = C.c_double()
    if lib.ref_time_scene(C.byref(d), warmup, iterations, C.byref(out)) != 0:
        raise RuntimeError("reference timing failed")
    return out.value
