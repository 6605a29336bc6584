"""ctypes binding of libdfpsr_hip.so (include/dfpsr_hip.h).

Loading fails loudly when the library is not built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# DFPSR_HIP_LIBRARY names another build of the library (the host-sanitized one, for instance).
LIB_PATH = os.environ.get("DFPSR_HIP_LIBRARY") or os.path.join(HERE, "libdfpsr_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "dfpsr_hip.h")

OK = 0
ERROR_ARGUMENT = -1
ERROR_STATE = -2
ERROR_DEVICE = -3
ERROR_MEMORY = -4


class DfpsrError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{message} (code {code})")
        self.code = code


class Image(C.Structure):
    _fields_ = [("pixels", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32),
                ("stride_bytes", C.c_int32), ("on_device", C.c_int32)]


class Camera(C.Structure):
    _fields_ = [("perspective", C.c_int32), ("location", C.c_float * 12), ("image_width", C.c_float),
                ("image_height", C.c_float), ("width_slope", C.c_float), ("near_clip", C.c_float),
                ("far_clip", C.c_float)]


class Tuning(C.Structure):
    _fields_ = [("tile_config", C.c_int32), ("persistent_per_cu", C.c_int32), ("use_graph", C.c_int32),
                ("clear", C.c_int32), ("clear_color", C.c_uint32), ("clear_depth", C.c_float),
                ("raster_path", C.c_int32), ("binning", C.c_int32)]


class Job(C.Structure):
    _fields_ = [("color", Image), ("depth", Image), ("model", C.c_void_p), ("transform", C.c_float * 12),
                ("camera", Camera)]


class View(C.Structure):
    _fields_ = [("color", C.c_void_p), ("visibility", C.c_void_p), ("depth", C.c_void_p), ("grad_color", C.c_void_p),
                ("width", C.c_int32), ("height", C.c_int32), ("stride_pixels", C.c_int32),
                ("transform", C.c_float * 12), ("camera", Camera)]


_lib = None

_P = C.c_void_p
_I = C.c_int32

_SIGNATURES = {
    "dfpsr_hip_last_error": ([], C.c_char_p),
    "dfpsr_hip_version": ([], C.c_char_p),
    "dfpsr_hip_texture_create": ([C.POINTER(_P), _I, _I, _I], C.c_int),
    "dfpsr_hip_texture_write_level0": ([_P, _P, _I], C.c_int),
    "dfpsr_hip_texture_generate_pyramid": ([_P], C.c_int),
    "dfpsr_hip_texture_info": ([_P, _P, _P, _P, _P], C.c_int),
    "dfpsr_hip_texture_read": ([_P, _P], C.c_int),
    "dfpsr_hip_texture_device_pointer": ([_P, C.POINTER(_P)], C.c_int),
    "dfpsr_hip_texture_destroy": ([_P], None),
    "dfpsr_hip_model_create": ([C.POINTER(_P), _I], C.c_int),
    "dfpsr_hip_model_set_points": ([_P, _I, _P], C.c_int),
    "dfpsr_hip_model_add_part": ([_P, _I, _P, _P, _P, _P, _P], C.c_int),
    "dfpsr_hip_model_bound": ([_P, _P, _P], C.c_int),
    "dfpsr_hip_model_import_dmf1": ([C.POINTER(_P), C.c_char_p, C.c_int64, _I, _I, _P, _P], C.c_int),
    "dfpsr_hip_model_get_points": ([_P, _P, _P], C.c_int),
    "dfpsr_hip_model_part_count": ([_P, _P], C.c_int),
    "dfpsr_hip_model_get_part": ([_P, _I, _P, _P, _P, _P, _P, _P], C.c_int),
    "dfpsr_hip_model_part_name": ([_P, _I, _P, _I, _P], C.c_int),
    "dfpsr_hip_model_render_depth": ([_P, _P, _P, _P], C.c_int),
    "dfpsr_hip_model_render_depth_noalloc": ([_P, _P, _P, _P], C.c_int),
    "dfpsr_hip_model_destroy": ([_P], None),
    "dfpsr_hip_renderer_create": ([C.POINTER(_P)], C.c_int),
    "dfpsr_hip_renderer_destroy": ([_P], None),
    "dfpsr_hip_renderer_begin": ([_P, C.POINTER(Image), C.POINTER(Image)], C.c_int),
    "dfpsr_hip_renderer_give_model": ([_P, _P, _P, C.POINTER(Camera)], C.c_int),
    "dfpsr_hip_renderer_give_triangles": ([_P, _I, _P, _P, _P, _P, _P, _I, C.POINTER(Camera)], C.c_int),
    "dfpsr_hip_renderer_give_projected_triangles": ([_P, _I, _P, _P, _P, _P, _P, _I, C.POINTER(Camera)], C.c_int),
    "dfpsr_hip_renderer_occlude_from_box": ([_P, _P, _P, _P, C.POINTER(Camera)], C.c_int),
    "dfpsr_hip_renderer_occlude_from_top_rows": ([_P, C.POINTER(Camera)], C.c_int),
    "dfpsr_hip_renderer_occlude_from_existing_triangles": ([_P], C.c_int),
    "dfpsr_hip_renderer_is_box_visible": ([_P, _P, _P, _P, C.POINTER(Camera), _P], C.c_int),
    "dfpsr_hip_renderer_has_occluders": ([_P, _P], C.c_int),
    "dfpsr_hip_renderer_occlusion_grid": ([_P, _P, _P, _P], C.c_int),
    "dfpsr_hip_renderer_takes_triangles": ([_P, _P], C.c_int),
    "dfpsr_hip_renderer_end": ([_P], C.c_int),
    "dfpsr_hip_renderer_stats": ([_P, _P], C.c_int),
    "dfpsr_hip_renderer_set_timing": ([_P, _I], C.c_int),
    "dfpsr_hip_renderer_timing": ([_P, _P], C.c_int),
    "dfpsr_hip_renderer_set_tuning": ([_P, C.POINTER(Tuning)], C.c_int),
    "dfpsr_hip_batch_create": ([C.POINTER(_P)], C.c_int),
    "dfpsr_hip_batch_destroy": ([_P], None),
    "dfpsr_hip_batch_render": ([_P, C.POINTER(Job), _I, C.POINTER(Tuning)], C.c_int),
    "dfpsr_hip_batch_stats": ([_P, _P], C.c_int),
    "dfpsr_hip_batch_set_timing": ([_P, _I], C.c_int),
    "dfpsr_hip_batch_timing": ([_P, _P], C.c_int),
    "dfpsr_hip_batch_autotune": ([_P, C.POINTER(Job), _I, _I, C.POINTER(Tuning), C.POINTER(Tuning), _P], C.c_int),
    "dfpsr_hip_diff_create": ([C.POINTER(_P)], C.c_int),
    "dfpsr_hip_diff_destroy": ([_P], None),
    "dfpsr_hip_diff_forward": ([_P, _P, C.POINTER(View), _I, _P, _I, _I, _P, C.POINTER(Tuning)], C.c_int),
    "dfpsr_hip_diff_backward": ([_P, _P, _P, C.POINTER(Tuning)], C.c_int),
    "dfpsr_hip_diff_stats": ([_P, _P], C.c_int),
    "dfpsr_hip_diff_set_timing": ([_P, _I], C.c_int),
    "dfpsr_hip_diff_timing": ([_P, _P], C.c_int),
    "dfpsr_hip_probe_edge_side": ([_P, _P, _P, _P, _I, _P], C.c_int),
}


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built (run python -m dfpsr_amd.build)")
        l = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(l, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = res
        _lib = l
    return _lib


def check(code: int) -> None:
    if code != OK:
        raise DfpsrError(code, lib().dfpsr_hip_last_error().decode())


def declared_symbols() -> list:
    """Every function name declared in include/dfpsr_hip.h."""
    import re
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"\b(dfpsr_hip_[a-z0-9_]+)\s*\(", text)))
