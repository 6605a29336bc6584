"""Binary scene files for tests/host/facade_render (the C++14 facade's test program).

Layout, little endian: magic 'DFS1', int32 width, height, filter, perspective, point count,
part count, polygon count, texture count, instance count; uint32 initial colour; float32
initial depth, camera location[12], width slope, near clip, far clip; points (P x 3 float32); per texture int32 width, height,
resolutions and level 0 (power-of-two size, uint32); per part int32 polygon start, end, diffuse,
light; indices (M x 4 int32); colours (M x 4 x 4 float32); texture coordinates
(M x 4 x 4 float32); instance transforms (I x 12 float32).
"""
import numpy as np


def _pow2_level0(t):
    w = 1 << max(0, int(t.width - 1).bit_length())
    h = 1 << max(0, int(t.height - 1).bit_length())
    ys = np.arange(h) % t.height
    xs = np.arange(w) % t.width
    return w, h, np.ascontiguousarray(t.level0[ys][:, xs], dtype=np.uint32)


def _uniform(image, default):
    if image is None:
        return default
    flat = np.asarray(image).reshape(-1)
    assert np.all(flat.view(np.uint32) == flat[:1].view(np.uint32)), "scene files hold uniform initial canvases only"
    return flat[0]


def write_scene(scene, path):
    parts = []
    with open(path, "wb") as f:
        f.write(np.array([0x31534644], np.uint32).tobytes())
        f.write(np.array([scene.width, scene.height, scene.filter, int(scene.perspective), scene.points.shape[0],
                          len(scene.parts), scene.indices.shape[0], len(scene.textures), scene.instances.shape[0]],
                         np.int32).tobytes())
        f.write(np.array([_uniform(scene.initial_color, 0)], np.uint32).tobytes())
        f.write(np.array([_uniform(scene.initial_depth, 0.0)], np.float32).tobytes())
        f.write(np.asarray(scene.cam_location, np.float32).tobytes())
        f.write(np.array([scene.width_slope, scene.near_clip, scene.far_clip], np.float32).tobytes())
        f.write(np.ascontiguousarray(scene.points, np.float32).tobytes())
        for t in scene.textures:
            w, h, level0 = _pow2_level0(t)
            f.write(np.array([w, h, t.resolutions], np.int32).tobytes())
            f.write(level0.tobytes())
        for p in scene.parts:
            parts.append([p.polygon_start, p.polygon_end, p.diffuse, p.light])
        f.write(np.array(parts, np.int32).reshape(-1).tobytes())
        f.write(np.ascontiguousarray(scene.indices, np.int32).tobytes())
        f.write(np.ascontiguousarray(scene.colors, np.float32).tobytes())
        f.write(np.ascontiguousarray(scene.texcoords, np.float32).tobytes())
        f.write(np.ascontiguousarray(scene.instances, np.float32).tobytes())


def read_output(path, width, height):
    raw = np.fromfile(path, np.uint8)
    n = width * height
    color = raw[:n * 4].view(np.uint32).reshape(height, width)
    depth = raw[n * 4:n * 8].view(np.float32).reshape(height, width)
    return color, depth
