"""Python mirror of the reference's rendererAPI / modelAPI / textureAPI for the polygon path,
over the C ABI of libdfpsr_hip.so.  Names and argument meaning follow the reference
(api/rendererAPI.h, api/modelAPI.h, api/textureAPI.h); errors raise DfpsrError where the
reference calls throwError.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from ._lib import Camera, DfpsrError, Image, Job, Tuning, View, check, lib

Filter_Solid = 0
Filter_Alpha = 1


class Texture:
    """TextureRgbaU8 resident on the device."""

    def __init__(self, handle):
        self._h = handle
        self._destroy = lib().dfpsr_hip_texture_destroy

    def __del__(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)  # held by the object: module globals may be gone at exit
            self._h = None

    @property
    def handle(self):
        return self._h

    def info(self):
        vals = [C.c_int32() for _ in range(4)]
        check(lib().dfpsr_hip_texture_info(self._h, *[C.byref(v) for v in vals]))
        return tuple(v.value for v in vals)  # log2w, log2h, max_mip, texel_count

    def read(self) -> np.ndarray:
        n = self.info()[3]
        out = np.zeros(n, dtype=np.uint32)
        check(lib().dfpsr_hip_texture_read(self._h, out.ctypes.data_as(C.c_void_p)))
        return out


def texture_create_RgbaU8(width: int, height: int, resolutions: int) -> Texture:
    h = C.c_void_p()
    check(lib().dfpsr_hip_texture_create(C.byref(h), width, height, resolutions))
    return Texture(h)


def texture_write_level0(texture: Texture, pixels: np.ndarray) -> None:
    p = np.ascontiguousarray(pixels, dtype=np.uint32)
    check(lib().dfpsr_hip_texture_write_level0(texture.handle, p.ctypes.data_as(C.c_void_p), p.shape[1]))


def texture_generatePyramid(texture: Texture) -> None:
    check(lib().dfpsr_hip_texture_generate_pyramid(texture.handle))


class Model:
    def __init__(self, handle):
        self._h = handle
        self._destroy = lib().dfpsr_hip_model_destroy
        self._keep = []  # textures referenced by parts must outlive the model's frames

    def __del__(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)  # held by the object: module globals may be gone at exit
            self._h = None

    @property
    def handle(self):
        return self._h


def model_create(filter: int = Filter_Solid) -> Model:
    h = C.c_void_p()
    check(lib().dfpsr_hip_model_create(C.byref(h), filter))
    return Model(h)


def model_setPoints(model: Model, points: np.ndarray) -> None:
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    check(lib().dfpsr_hip_model_set_points(model.handle, p.shape[0], p.ctypes.data_as(C.c_void_p)))


def model_addPart(model: Model, indices, colors, texcoords, diffuse: Optional[Texture] = None,
                  light: Optional[Texture] = None) -> None:
    ix = np.ascontiguousarray(indices, dtype=np.int32).reshape(-1, 4)
    co = np.ascontiguousarray(colors, dtype=np.float32).reshape(-1, 16)
    tc = np.ascontiguousarray(texcoords, dtype=np.float32).reshape(-1, 16)
    check(lib().dfpsr_hip_model_add_part(model.handle, ix.shape[0], ix.ctypes.data_as(C.c_void_p),
                                         co.ctypes.data_as(C.c_void_p), tc.ctypes.data_as(C.c_void_p),
                                         diffuse.handle if diffuse else None, light.handle if light else None))
    model._keep += [t for t in (diffuse, light) if t is not None]


def model_importDMF1(content: bytes, detail_level: int = 2, pool: Optional[dict] = None) -> Model:
    """importFromContent_DMF1 (modelAPI.h:319): `pool` maps texture names to Textures (the
    caller's resource pool; names compared ignoring ASCII case)."""
    pool = pool or {}
    names = (C.c_char_p * max(1, len(pool)))(*[n.encode() for n in pool])
    handles = (C.c_void_p * max(1, len(pool)))(*[t.handle.value if hasattr(t.handle, "value") else t.handle for t in pool.values()])
    h = C.c_void_p()
    check(lib().dfpsr_hip_model_import_dmf1(C.byref(h), content, len(content), detail_level, len(pool),
                                            C.cast(names, C.c_void_p), C.cast(handles, C.c_void_p)))
    m = Model(h)
    m._keep += list(pool.values())
    return m


def model_getPoints(model: Model) -> np.ndarray:
    n = C.c_int32()
    check(lib().dfpsr_hip_model_get_points(model.handle, C.byref(n), None))
    out = np.zeros((n.value, 3), np.float32)
    check(lib().dfpsr_hip_model_get_points(model.handle, C.byref(n), out.ctypes.data_as(C.c_void_p)))
    return out


def model_getNumberOfParts(model: Model) -> int:
    n = C.c_int32()
    check(lib().dfpsr_hip_model_part_count(model.handle, C.byref(n)))
    return n.value


def model_getPart(model: Model, part: int) -> dict:
    """One part: indices (N, 4), colors (N, 4, 4), texcoords (N, 4, 4) and the raw handles of
    its diffuse and light textures (None when absent)."""
    n = C.c_int32()
    check(lib().dfpsr_hip_model_get_part(model.handle, part, C.byref(n), None, None, None, None, None))
    idx = np.zeros((n.value, 4), np.int32)
    col = np.zeros((n.value, 4, 4), np.float32)
    tex = np.zeros((n.value, 4, 4), np.float32)
    d, l = C.c_void_p(), C.c_void_p()
    check(lib().dfpsr_hip_model_get_part(model.handle, part, C.byref(n), idx.ctypes.data_as(C.c_void_p),
                                         col.ctypes.data_as(C.c_void_p), tex.ctypes.data_as(C.c_void_p), C.byref(d), C.byref(l)))
    return {"indices": idx, "colors": col, "texcoords": tex, "diffuse": d.value, "light": l.value}


def model_getPartName(model: Model, part: int) -> str:
    n = C.c_int32()
    check(lib().dfpsr_hip_model_part_name(model.handle, part, None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    check(lib().dfpsr_hip_model_part_name(model.handle, part, buf, n.value + 1, C.byref(n)))
    return buf.value.decode("utf-8", "replace")


def model_getBoundingBox(model: Model):
    mn = np.zeros(3, np.float32)
    mx = np.zeros(3, np.float32)
    check(lib().dfpsr_hip_model_bound(model.handle, mn.ctypes.data_as(C.c_void_p), mx.ctypes.data_as(C.c_void_p)))
    return mn, mx


def model_renderDepth(model: Model, transform, depth: Image, camera: Camera) -> None:
    """model_renderDepth (modelAPI.h:291): the model's depth alone into `depth`, synchronously."""
    t = np.ascontiguousarray(transform, dtype=np.float32)
    check(lib().dfpsr_hip_model_render_depth(model.handle, t.ctypes.data_as(C.c_void_p), C.byref(depth), C.byref(camera)))


def camera_create(perspective: bool, location, image_width: float, image_height: float, width_slope: float,
                  near_clip: float = 0.01, far_clip: float = 1000.0) -> Camera:
    """Camera::createPerspective (perspective=True) or Camera::createOrthogonal."""
    c = Camera()
    c.perspective = int(perspective)
    c.location[:] = [float(v) for v in np.asarray(location, dtype=np.float32)]
    c.image_width, c.image_height = float(image_width), float(image_height)
    c.width_slope, c.near_clip, c.far_clip = float(width_slope), float(near_clip), float(far_clip)
    return c


class Renderer:
    def __init__(self, handle):
        self._h = handle
        self._destroy = lib().dfpsr_hip_renderer_destroy
        self._images = None

    def __del__(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)  # held by the object: module globals may be gone at exit
            self._h = None

    @property
    def handle(self):
        return self._h

    def set_timing(self, enable: bool) -> None:
        check(lib().dfpsr_hip_renderer_set_timing(self._h, int(enable)))

    def timing(self):
        out = np.zeros(5, np.float32)
        check(lib().dfpsr_hip_renderer_timing(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def stats(self):
        out = np.zeros(4, np.int64)
        check(lib().dfpsr_hip_renderer_stats(self._h, out.ctypes.data_as(C.c_void_p)))
        return out


def renderer_create() -> Renderer:
    h = C.c_void_p()
    check(lib().dfpsr_hip_renderer_create(C.byref(h)))
    return Renderer(h)


def host_image(array: Optional[np.ndarray]) -> Image:
    """Canvas over a host numpy array (H, W) of uint32 or float32, copied at begin/end."""
    im = Image()
    if array is None:
        return im
    assert array.flags["C_CONTIGUOUS"] and array.itemsize == 4
    im.pixels = array.ctypes.data
    im.height, im.width = array.shape
    im.stride_bytes = array.strides[0]
    im.on_device = 0
    return im


def device_image(tensor) -> Image:
    """Canvas over a device tensor (H, W) of 4-byte elements (torch, on the GPU), drawn in place."""
    im = Image()
    if tensor is None:
        return im
    im.pixels = tensor.data_ptr()
    im.height, im.width = tensor.shape
    im.stride_bytes = tensor.stride(0) * tensor.element_size()
    im.on_device = 1
    return im


def renderer_begin(renderer: Renderer, color: Image, depth: Image) -> None:
    renderer._images = (color, depth)
    check(lib().dfpsr_hip_renderer_begin(renderer.handle, C.byref(color), C.byref(depth)))


def renderer_giveTask(renderer: Renderer, model: Model, transform, camera: Camera) -> None:
    t = np.ascontiguousarray(transform, dtype=np.float32)
    check(lib().dfpsr_hip_renderer_give_model(renderer.handle, model.handle, t.ctypes.data_as(C.c_void_p), C.byref(camera)))


def renderer_giveTask_triangles(renderer: Renderer, camera_space, colors, texcoords, diffuse: Optional[Texture],
                                light: Optional[Texture], filter: int, camera: Camera) -> None:
    cs = np.ascontiguousarray(camera_space, dtype=np.float32).reshape(-1, 9)
    co = np.ascontiguousarray(colors, dtype=np.float32).reshape(-1, 12)
    tc = np.ascontiguousarray(texcoords, dtype=np.float32).reshape(-1, 12)
    check(lib().dfpsr_hip_renderer_give_triangles(renderer.handle, cs.shape[0], cs.ctypes.data_as(C.c_void_p),
                                                  co.ctypes.data_as(C.c_void_p), tc.ctypes.data_as(C.c_void_p),
                                                  diffuse.handle if diffuse else None, light.handle if light else None,
                                                  filter, C.byref(camera)))


PROJECTED_DTYPE = np.dtype({"names": ["cs", "is", "flat"], "formats": [(np.float32, 3), (np.float32, 2), (np.int64, 2)],
                            "offsets": [0, 12, 24], "itemsize": 40})


def renderer_giveTask_projected(renderer: Renderer, corners, colors, texcoords, diffuse: Optional[Texture],
                                light: Optional[Texture], filter: int, camera: Camera) -> None:
    """renderer_giveTask_triangle with the caller's ProjectedPoints: corners is an (N, 3) array
    of PROJECTED_DTYPE records (cs, is, flat)."""
    pp = np.ascontiguousarray(corners)
    assert pp.dtype == PROJECTED_DTYPE and pp.ndim == 2 and pp.shape[1] == 3
    co = np.ascontiguousarray(colors, dtype=np.float32).reshape(-1, 12)
    tc = np.ascontiguousarray(texcoords, dtype=np.float32).reshape(-1, 12)
    assert co.shape[0] == pp.shape[0] and tc.shape[0] == pp.shape[0]
    check(lib().dfpsr_hip_renderer_give_projected_triangles(renderer.handle, pp.shape[0], pp.ctypes.data_as(C.c_void_p),
                                                            co.ctypes.data_as(C.c_void_p), tc.ctypes.data_as(C.c_void_p),
                                                            diffuse.handle if diffuse else None,
                                                            light.handle if light else None, filter, C.byref(camera)))


def _vec3(v):
    return np.ascontiguousarray(v, dtype=np.float32).reshape(3)


def renderer_occludeFromBox(renderer: Renderer, minimum, maximum, transform, camera: Camera) -> None:
    mn, mx, t = _vec3(minimum), _vec3(maximum), np.ascontiguousarray(transform, dtype=np.float32)
    check(lib().dfpsr_hip_renderer_occlude_from_box(renderer.handle, mn.ctypes.data_as(C.c_void_p), mx.ctypes.data_as(C.c_void_p),
                                                    t.ctypes.data_as(C.c_void_p), C.byref(camera)))


def renderer_occludeFromTopRows(renderer: Renderer, camera: Camera) -> None:
    check(lib().dfpsr_hip_renderer_occlude_from_top_rows(renderer.handle, C.byref(camera)))


def renderer_occludeFromExistingTriangles(renderer: Renderer) -> None:
    check(lib().dfpsr_hip_renderer_occlude_from_existing_triangles(renderer.handle))


def renderer_isBoxVisible(renderer: Renderer, minimum, maximum, transform, camera: Camera) -> bool:
    mn, mx, t = _vec3(minimum), _vec3(maximum), np.ascontiguousarray(transform, dtype=np.float32)
    out = C.c_int32()
    check(lib().dfpsr_hip_renderer_is_box_visible(renderer.handle, mn.ctypes.data_as(C.c_void_p), mx.ctypes.data_as(C.c_void_p),
                                                  t.ctypes.data_as(C.c_void_p), C.byref(camera), C.byref(out)))
    return bool(out.value)


def renderer_hasOccluders(renderer: Renderer) -> bool:
    out = C.c_int32()
    check(lib().dfpsr_hip_renderer_has_occluders(renderer.handle, C.byref(out)))
    return bool(out.value)


def renderer_occlusionGrid(renderer: Renderer) -> np.ndarray:
    """The renderer's occlusion cells (16 x 16 pixels each) as they are now."""
    w, h = C.c_int32(), C.c_int32()
    check(lib().dfpsr_hip_renderer_occlusion_grid(renderer.handle, None, C.byref(w), C.byref(h)))
    out = np.zeros((h.value, w.value), np.float32)
    check(lib().dfpsr_hip_renderer_occlusion_grid(renderer.handle, out.ctypes.data_as(C.c_void_p), C.byref(w), C.byref(h)))
    return out


def renderer_takesTriangles(renderer: Renderer) -> bool:
    v = C.c_int32()
    check(lib().dfpsr_hip_renderer_takes_triangles(renderer.handle, C.byref(v)))
    return bool(v.value)


def renderer_end(renderer: Renderer) -> None:
    check(lib().dfpsr_hip_renderer_end(renderer.handle))


# ---------------------------------------------------------------------------------------
# Scenes (dfpsr_amd.scenes) on the GPU
# ---------------------------------------------------------------------------------------

class GpuScene:
    """A scene's textures and model uploaded once, ready for frames."""

    def __init__(self, scene):
        self.scene = scene
        self.textures = []
        for t in scene.textures:
            tex = texture_create_RgbaU8(t.width, t.height, t.resolutions)
            texture_write_level0(tex, _tile_level0(t))
            texture_generatePyramid(tex)
            self.textures.append(tex)
        self.model = model_create(scene.filter)
        model_setPoints(self.model, scene.points)
        for p in scene.parts:
            sl = slice(p.polygon_start, p.polygon_end)
            model_addPart(self.model, scene.indices[sl], scene.colors[sl], scene.texcoords[sl],
                          self.textures[p.diffuse] if p.diffuse >= 0 else None,
                          self.textures[p.light] if p.light >= 0 else None)
        self.camera = camera_create(scene.perspective, scene.cam_location, scene.width, scene.height, scene.width_slope,
                                    scene.near_clip, scene.far_clip)
        self._prepared = None

    def _prepared_arguments(self):
        """The instances' transforms and the camera as the C ABI takes them, made once and again
        only when the scene's instances or the camera object are replaced."""
        key = (id(self.scene.instances), id(self.camera))
        if self._prepared is None or self._prepared[0] != key:
            transforms = [np.ascontiguousarray(t, dtype=np.float32) for t in self.scene.instances]
            self._prepared = (key, transforms, [t.ctypes.data_as(C.c_void_p) for t in transforms], C.byref(self.camera),
                              (self.scene.instances, self.camera))
        return self._prepared

    def give(self, renderer: Renderer) -> None:
        sc = self.scene
        if sc.projected is None and not sc.occluders:
            # The common frame: one renderer_giveTask per instance with the prepared arguments.
            _, _, pointers, camera, _ = self._prepared_arguments()
            give_model = lib().dfpsr_hip_renderer_give_model
            for ptr in pointers:
                check(give_model(renderer.handle, self.model.handle, ptr, camera))
            return
        if sc.projected is not None:
            # One renderer_giveTask_triangle batch per part, in polygon order.
            for p in sc.parts:
                sl = slice(p.polygon_start, p.polygon_end)
                renderer_giveTask_projected(renderer, sc.projected[sl], sc.colors[sl, :3], sc.texcoords[sl, :3],
                                            self.textures[p.diffuse] if p.diffuse >= 0 else None,
                                            self.textures[p.light] if p.light >= 0 else None, sc.filter, self.camera)
            return
        ops = sc.occluders or []
        self.visible = []

        def occluders_before(i):
            for op in ops:
                if op[1] != i:
                    continue
                if op[0] == "box":
                    renderer_occludeFromBox(renderer, op[2], op[3], op[4], self.camera)
                elif op[0] == "top_rows":
                    renderer_occludeFromTopRows(renderer, self.camera)
                elif op[0] == "existing":
                    renderer_occludeFromExistingTriangles(renderer)
                elif op[0] == "query":
                    self.visible.append(renderer_isBoxVisible(renderer, op[2], op[3], op[4], self.camera))

        for i, t in enumerate(sc.instances):
            occluders_before(i)
            renderer_giveTask(renderer, self.model, t, self.camera)
        occluders_before(len(sc.instances))


def _tile_level0(t) -> np.ndarray:
    """Level 0 at the texture's power-of-two size, repeating the given pixels (as the reference
    driver fills texture_getSafePointer rows)."""
    w = 1 << max(0, int(t.width - 1).bit_length())
    h = 1 << max(0, int(t.height - 1).bit_length())
    if (w, h) == (t.width, t.height):
        return np.ascontiguousarray(t.level0, dtype=np.uint32)
    ys = np.arange(h) % t.height
    xs = np.arange(w) % t.width
    return np.ascontiguousarray(t.level0[ys][:, xs], dtype=np.uint32)


def render_scene(scene, renderer: Optional[Renderer] = None, gpu_scene: Optional[GpuScene] = None):
    """One frame of `scene` on host canvases; returns (color, depth) as numpy arrays."""
    r = renderer or renderer_create()
    g = gpu_scene or GpuScene(scene)
    color = np.zeros((scene.height, scene.width), np.uint32) if scene.initial_color is None else scene.initial_color.astype(np.uint32).copy()
    depth = np.zeros((scene.height, scene.width), np.float32) if scene.initial_depth is None else scene.initial_depth.astype(np.float32).copy()
    renderer_begin(r, host_image(color if scene.has_color else None), host_image(depth if scene.has_depth else None))
    g.give(r)
    renderer_end(r)
    return color, depth


def render_scene_depth(scene, gpu_scene: Optional[GpuScene] = None, on_device: bool = False):
    """model_renderDepth of every instance of `scene` into its depth canvas alone; returns it
    as a numpy array.  on_device: draw into a device tensor instead of a host array."""
    g = gpu_scene or GpuScene(scene)
    depth = np.zeros((scene.height, scene.width), np.float32) if scene.initial_depth is None else scene.initial_depth.astype(np.float32).copy()
    if on_device:
        import torch
        dev = torch.from_numpy(depth).cuda()
        for t in scene.instances:
            model_renderDepth(g.model, t, device_image(dev), g.camera)
        return dev.cpu().numpy()
    for t in scene.instances:
        model_renderDepth(g.model, t, host_image(depth), g.camera)
    return depth


# ---------------------------------------------------------------------------------------
# Tuning and batches
# ---------------------------------------------------------------------------------------

def tuning(tile_config: int = 0, persistent_per_cu: int = 0, use_graph: bool = False, clear: bool = False,
           clear_color: int = 0, clear_depth: float = 0.0, raster_path: int = 0, binning: int = 0) -> Tuning:
    t = Tuning()
    t.tile_config, t.persistent_per_cu, t.use_graph = tile_config, persistent_per_cu, int(use_graph)
    t.clear, t.clear_color, t.clear_depth = int(clear), clear_color, clear_depth
    t.raster_path = raster_path
    t.binning = binning
    return t


def renderer_setTuning(renderer: Renderer, t: Tuning) -> None:
    check(lib().dfpsr_hip_renderer_set_tuning(renderer.handle, C.byref(t)))


class Batch:
    def __init__(self):
        h = C.c_void_p()
        check(lib().dfpsr_hip_batch_create(C.byref(h)))
        self._h = h
        self._destroy = lib().dfpsr_hip_batch_destroy

    def __del__(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)  # held by the object: module globals may be gone at exit
            self._h = None

    def render(self, jobs, t: Tuning) -> None:
        check(lib().dfpsr_hip_batch_render(self._h, jobs, len(jobs), C.byref(t)))

    def set_timing(self, enable: bool) -> None:
        check(lib().dfpsr_hip_batch_set_timing(self._h, int(enable)))

    def timing(self):
        out = np.zeros(5, np.float32)
        check(lib().dfpsr_hip_batch_timing(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def stats(self):
        out = np.zeros(4, np.int64)
        check(lib().dfpsr_hip_batch_stats(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def autotune(self, jobs, repeats: int = 3, base: Optional[Tuning] = None):
        """Sweeps tile shape x persistent workgroups per CU (graph replay on) and returns
        (best Tuning, 3 x 5 array of median milliseconds)."""
        best = Tuning()
        ms = np.zeros((3, 5), np.float32)
        check(lib().dfpsr_hip_batch_autotune(self._h, jobs, len(jobs), repeats, C.byref(base) if base is not None else None,
                                             C.byref(best), ms.ctypes.data_as(C.c_void_p)))
        return best, ms


class GpuJobs:
    """The jobs of a batch resident on the GPU: shared textures, one model per job and device
    canvases (torch tensors), with the job array for dfpsr_hip_batch_render."""

    def __init__(self, scenes_list, texture_pool=None):
        import torch
        self.scenes = scenes_list
        self.textures = {}
        self.gpu = []
        self.colors, self.depths = [], []
        self.jobs = (Job * len(scenes_list))()
        for i, sc in enumerate(scenes_list):
            texs = []
            ids = getattr(sc, "texture_ids", None)
            for k, t in enumerate(sc.textures):
                key = ("shared", ids[k]) if ids else ("own", id(t))
                if key not in self.textures:
                    tex = texture_create_RgbaU8(t.width, t.height, t.resolutions)
                    texture_write_level0(tex, _tile_level0(t))
                    texture_generatePyramid(tex)
                    self.textures[key] = tex
                texs.append(self.textures[key])
            model = model_create(sc.filter)
            model_setPoints(model, sc.points)
            for p in sc.parts:
                sl = slice(p.polygon_start, p.polygon_end)
                model_addPart(model, sc.indices[sl], sc.colors[sl], sc.texcoords[sl],
                              texs[p.diffuse] if p.diffuse >= 0 else None, texs[p.light] if p.light >= 0 else None)
            color = torch.zeros((sc.height, sc.width), dtype=torch.int32, device="cuda")
            depth = torch.zeros((sc.height, sc.width), dtype=torch.float32, device="cuda")
            cam = camera_create(sc.perspective, sc.cam_location, sc.width, sc.height, sc.width_slope, sc.near_clip, sc.far_clip)
            j = self.jobs[i]
            j.color = device_image(color)
            j.depth = device_image(depth)
            j.model = model.handle
            j.transform[:] = [float(v) for v in np.asarray(sc.instances[0], dtype=np.float32)]
            j.camera = cam
            self.gpu.append(model)
            self.colors.append(color)
            self.depths.append(depth)

    def clear(self):
        for c, d in zip(self.colors, self.depths):
            c.zero_()
            d.zero_()

    def image(self, i):
        return self.colors[i].cpu().numpy().view(np.uint32), self.depths[i].cpu().numpy()

    def render_plain(self, renderer: Renderer) -> None:
        """The per-job loop of the plain path: renderer_begin / giveTask / end per job."""
        for i, j in enumerate(self.jobs):
            renderer_begin(renderer, j.color, j.depth)
            check(lib().dfpsr_hip_renderer_give_model(renderer.handle, j.model, j.transform, C.byref(j.camera)))
            renderer_end(renderer)


# ---------------------------------------------------------------------------------------
# Differentiable rendering
# ---------------------------------------------------------------------------------------

class DiffRenderer:
    """dfpsr_hip_diff: forward with the differentiable float shader, adjoint with visibility
    fixed from the forward.  Buffers are torch tensors on the GPU."""

    def __init__(self):
        h = C.c_void_p()
        check(lib().dfpsr_hip_diff_create(C.byref(h)))
        self._h = h
        self._destroy = lib().dfpsr_hip_diff_destroy

    def __del__(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)  # held by the object: module globals may be gone at exit
            self._h = None

    def forward(self, model: Model, views, texture, multipliers=None, t: Optional[Tuning] = None) -> None:
        import math
        t = t or tuning(clear=True)
        h, w = texture.shape[0], texture.shape[1]
        assert texture.is_contiguous() and texture.shape[2] == 4
        check(lib().dfpsr_hip_diff_forward(self._h, model.handle, views, len(views), texture.data_ptr(),
                                           int(math.log2(w)), int(math.log2(h)),
                                           multipliers.data_ptr() if multipliers is not None else None, C.byref(t)))

    def backward(self, grad_texture=None, grad_multipliers=None, t: Optional[Tuning] = None) -> None:
        t = t or tuning()
        check(lib().dfpsr_hip_diff_backward(self._h, grad_texture.data_ptr() if grad_texture is not None else None,
                                            grad_multipliers.data_ptr() if grad_multipliers is not None else None,
                                            C.byref(t)))

    def set_timing(self, enable: bool) -> None:
        check(lib().dfpsr_hip_diff_set_timing(self._h, int(enable)))

    def timing(self):
        out = np.zeros(5, np.float32)
        check(lib().dfpsr_hip_diff_timing(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def stats(self):
        out = np.zeros(4, np.int64)
        check(lib().dfpsr_hip_diff_stats(self._h, out.ctypes.data_as(C.c_void_p)))
        return out


class GpuViews:
    """Views of one model for the differentiable renderer: the model (points, triangles, uv)
    and per view colour / visibility / depth / gradient buffers as torch tensors."""

    def __init__(self, view_scenes):
        import torch
        first = view_scenes[0]
        self.model = model_create(0)
        model_setPoints(self.model, first.points)
        model_addPart(self.model, first.indices, first.colors, first.texcoords)
        self.point_count = first.points.shape[0]
        n = len(view_scenes)
        self.views = (View * n)()
        self.color, self.visibility, self.depth, self.grad = [], [], [], []
        for i, sc in enumerate(view_scenes):
            col = torch.zeros((sc.height, sc.width, 4), dtype=torch.float32, device="cuda")
            vis = torch.full((sc.height, sc.width), -1, dtype=torch.int32, device="cuda")
            dep = torch.zeros((sc.height, sc.width), dtype=torch.float32, device="cuda")
            grad = torch.zeros((sc.height, sc.width, 4), dtype=torch.float32, device="cuda")
            v = self.views[i]
            v.color, v.visibility, v.depth, v.grad_color = col.data_ptr(), vis.data_ptr(), dep.data_ptr(), grad.data_ptr()
            v.width, v.height, v.stride_pixels = sc.width, sc.height, sc.width
            v.transform[:] = [float(x) for x in np.asarray(sc.instances[0], dtype=np.float32)]
            v.camera = camera_create(sc.perspective, sc.cam_location, sc.width, sc.height, sc.width_slope,
                                     sc.near_clip, sc.far_clip)
            self.color.append(col)
            self.visibility.append(vis)
            self.depth.append(dep)
            self.grad.append(grad)


__all__ = [n for n in dir() if not n.startswith("_")] + ["DfpsrError"]
