"""Writes the golden fixtures under tests/golden/ from the reference renderer.

Run in the build container only (it needs /root/reference and oracle/_ref/libdfpsr_ref.so,
built by `make -f oracle/Makefile.ref`):

    python tests/golden/make_goldens.py

Fixtures written (all data, no reference source):
  known_answers.json    known-answer vectors taken from the reference's own unit tests
                        (test/tests/TextureTest.cpp, test/tests/SimdTest.cpp), extracted as data;
  scenes.npz            colour + depth of every scene in tests/fixture_scenes.py, and the answers of
                        the renderer_isBoxVisible calls of scenes with occluder calls;
  config0.npz           colour + depth of the 640x480 cube-soup frame;
  config1.json          SHA-256 of the 1920x1080 terrain frame (colour, depth, id-colour) + crops;
  config1_crops.npz     64x64 crops of that frame;
  texture_vectors.npz   a 64x64 pyramid and mip-selected bilinear samples of random quads;
  raster_vectors.npz    row intervals of random small triangles;
  reciprocal.npz        the SSE reciprocal-with-Newton-step of sample values.
Colour outputs go through the CPU's rcpps estimate and are tagged with the CPU vendor.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import harness  # noqa: E402

REF = harness.load_ref()  # load before any large numpy work (its static heap set-up needs it)

from dfpsr_amd import scenes as S  # noqa: E402
import fixture_scenes  # noqa: E402

REFERENCE_TESTS = "/root/reference/Source/test/tests"


def cpu_vendor() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("vendor_id"):
                return line.split(":")[1].strip()
    except OSError:
        pass
    return "unknown"


def f32(expr: str) -> float:
    """Evaluate a float literal expression such as '0.5f / 4.0f' or '0.01f - 256.0f' in float32."""
    expr = expr.strip()
    m = re.fullmatch(r"(-?[0-9.]+)f\s*([-+*/])\s*(-?[0-9.]+)f", expr)
    if m:
        a, op, b = np.float32(m.group(1)), m.group(2), np.float32(m.group(3))
        return float({"+": a + b, "-": a - b, "*": a * b, "/": a / b}[op])
    return float(np.float32(expr.rstrip("f")))


def num(s: str) -> int:
    s = s.strip().rstrip("u")
    return int(s, 2) if s.startswith("0b") else int(s)


def known_answers() -> dict:
    """Texture layout, read, nearest-sample and linear-mix known answers from TextureTest.cpp;
    reciprocal and clampUpper lanes from SimdTest.cpp."""
    text = open(os.path.join(REFERENCE_TESTS, "TextureTest.cpp"), encoding="utf-8-sig").read()
    blocks = []
    current = None
    for line in text.splitlines():
        m = re.search(r"TextureRgbaU8 texture = TextureRgbaU8\((\d+),\s*(\d+)(?:,\s*(\d+))?\)", line)
        if m:
            current = {"log2w": int(m.group(1)), "log2h": int(m.group(2)),
                       "max_mip": int(m.group(3)) if m.group(3) else 15, "writes": [], "checks": []}
            blocks.append(current)
            continue
        if current is None:
            continue
        m = re.search(r"texture_writePixel\(texture,\s*([^,]+),\s*([^,]+),\s*([^,]+),\s*([^)]+)\)", line)
        if m:
            current["writes"].append([num(m.group(i)) for i in range(1, 5)])
            continue
        m = re.search(r"ASSERT_EQUAL\(texture\.impl_startOffset\s*,\s*(\S+)\)", line)
        if m:
            current["checks"].append({"op": 0, "expected": num(m.group(1))})
            continue
        m = re.search(r"ASSERT_EQUAL\(texture\.impl_maxLevelMask\s*,\s*(\S+)\)", line)
        if m:
            current["checks"].append({"op": 1, "expected": num(m.group(1))})
            continue
        m = re.search(r"ASSERT_EQUAL\(texture_getSmallestMipLevel\(texture\),\s*(\d+)\)", line)
        if m:
            current["checks"].append({"op": 6, "expected": int(m.group(1))})
            continue
        m = re.search(r"ASSERT_EQUAL\(texture_getPixelOffsetToLayer\(texture,\s*(\w+)\),\s*(\w+)\)", line)
        if m:
            current["checks"].append({"op": 2, "mip": num(m.group(1)), "expected": num(m.group(2))})
            continue
        m = re.search(r"ASSERT_EQUAL\(texture_getPixelOffset\(texture,\s*(\w+),\s*(\w+),\s*(\w+)\),\s*(\w+)\)", line)
        if m:
            current["checks"].append({"op": 3, "x": num(m.group(1)), "y": num(m.group(2)), "mip": num(m.group(3)),
                                      "expected": num(m.group(4))})
            continue
        m = re.search(r"ASSERT_EQUAL\(texture_readPixel\(texture,\s*(\w+),\s*(\w+),\s*(\w+)\),\s*(\w+)\)", line)
        if m:
            current["checks"].append({"op": 4, "x": num(m.group(1)), "y": num(m.group(2)), "mip": num(m.group(3)),
                                      "expected": num(m.group(4))})
            continue
        m = re.search(r"ASSERT_EQUAL\(texture_sample_nearest\(texture,\s*([^,]+),\s*([^,]+),\s*(\w+)\),\s*(\w+)\)", line)
        if m:
            current["checks"].append({"op": 5, "u": f32(m.group(1)), "v": f32(m.group(2)), "mip": num(m.group(3)),
                                      "expected": num(m.group(4))})
            continue

    # The linear mix case: four lanes of packed colours A, B, weights and the expected result.
    lin = re.search(r"texture_interpolate_color_linear<U32x4>\((.*?)\);\s*U32x4 expectedColor = (.*?);", text, re.S)
    lanes = lambda s: [[int(v) for v in g.split(",")] for g in re.findall(r"U32x4\(\s*([0-9,\s]+)\)", s)]
    args = lanes(lin.group(1))
    expected = lanes(lin.group(2))

    def pack(ch):  # four channel vectors -> four packed lanes
        return [ch[0][i] | (ch[1][i] << 8) | (ch[2][i] << 16) | (ch[3][i] << 24) for i in range(4)]

    linear = {"a": pack(args[0:4]), "b": pack(args[4:8]), "weight": args[8], "expected": pack(expected)}

    simd = open(os.path.join(REFERENCE_TESTS, "SimdTest.cpp"), encoding="utf-8-sig").read()
    rec = re.search(r"reciprocal\(F32x4\(([^)]*)\)\), F32x4\(([^)]*)\)", simd)
    recip = {"in": [f32(v) for v in rec.group(1).split(",")], "expected": [f32(v) for v in rec.group(2).split(",")]}
    cu = re.search(r"clampUpper\(F32x4\(([^)]*)\), F32x4\(([^)]*)\)\), F32x4\(([^)]*)\)", simd)
    clamp = {"in": [f32(v) for v in cu.group(1).split(",")], "limit": f32(cu.group(2)),
             "expected": [f32(v) for v in cu.group(3).split(",")]}
    return {"source": ["test/tests/TextureTest.cpp", "test/tests/SimdTest.cpp"], "textures": blocks,
            "linear_mix": linear, "reciprocal": recip, "clamp_upper": clamp}


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    vendor = cpu_vendor()
    ka = known_answers()
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(ka, f, indent=1)
    n_checks = sum(len(b["checks"]) for b in ka["textures"])
    print(f"known answers: {len(ka['textures'])} texture blocks, {n_checks} checks")

    # Scene fixtures.
    data = {}
    names = []
    for sc in fixture_scenes.fixture_scenes():
        names.append(sc.name)
        if sc.occluders:
            c, d, visible = harness.ref_render_queries(sc)
            data[sc.name + "/visible"] = visible
        else:
            c, d = harness.ref_render(sc, True)
        data[sc.name + "/color"] = c
        data[sc.name + "/depth"] = d
    np.savez_compressed(os.path.join(HERE, "scenes.npz"), names=np.array(names), vendor=np.array(vendor), **data)
    print(f"scenes: {len(names)}")

    c0 = S.cube_soup()
    c, d = harness.ref_render(c0, True)
    cid, did = harness.ref_render(fixture_scenes.id_color(c0), True)
    np.savez_compressed(os.path.join(HERE, "config0.npz"), color=c, depth=d, id_color=cid, vendor=np.array(vendor))

    c1 = S.terrain()
    c, d = harness.ref_render(c1, True)
    cid, did = harness.ref_render(fixture_scenes.id_color(c1), True)
    crops = {}
    for i, (y, x) in enumerate([(0, 0), (500, 900), (1016, 1856), (300, 64), (800, 1500)]):
        crops[f"crop{i}_color"] = c[y:y + 64, x:x + 64]
        crops[f"crop{i}_depth"] = d[y:y + 64, x:x + 64]
        crops[f"crop{i}_origin"] = np.array([y, x])
    np.savez_compressed(os.path.join(HERE, "config1_crops.npz"), **crops)
    with open(os.path.join(HERE, "config1.json"), "w") as f:
        json.dump({"vendor": vendor, "color_sha256": sha(c), "depth_sha256": sha(d), "id_color_sha256": sha(cid),
                   "id_depth_sha256": sha(did), "covered": float(np.mean(d > 0))}, f, indent=1)

    # Texture vectors: pyramid and mip-selected bilinear quads (integer results).
    tex = S.noise_checker_texture(64, 4242, 7, check=8)
    cap = 64 * 64 * 2
    buf = np.zeros(cap, dtype=np.uint32)
    l2w, l2h, mm = C.c_int32(), C.c_int32(), C.c_int32()
    REF.ref_texture_pyramid.restype = C.c_int32
    total = REF.ref_texture_pyramid(64, 64, 7, tex.level0.ctypes.data_as(C.c_void_p), buf.ctypes.data_as(C.c_void_p), cap,
                                    C.byref(l2w), C.byref(l2h), C.byref(mm))
    rng = S.Rng(99)
    q = 4096
    base_u = rng.uniform(q, -3.0, 4.0)
    base_v = rng.uniform(q, -3.0, 4.0)
    du = rng.uniform(q, -0.3, 0.3) * rng.uniform(q, 0.0, 1.0) ** 4
    dv = rng.uniform(q, -0.3, 0.3) * rng.uniform(q, 0.0, 1.0) ** 4
    u = np.stack([base_u, base_u + du, base_u + dv * 0.5, base_u + du + dv * 0.5], 1).astype(np.float32)
    v = np.stack([base_v, base_v + dv * 0.3, base_v + dv, base_v + dv * 1.3], 1).astype(np.float32)
    colors = np.zeros((q, 4), dtype=np.uint32)
    mips = np.zeros(q, dtype=np.uint32)
    REF.ref_sample_quads(64, 64, 7, tex.level0.ctypes.data_as(C.c_void_p), q, u.ctypes.data_as(C.c_void_p),
                         v.ctypes.data_as(C.c_void_p), colors.ctypes.data_as(C.c_void_p), mips.ctypes.data_as(C.c_void_p))
    np.savez_compressed(os.path.join(HERE, "texture_vectors.npz"), level0=tex.level0, pyramid=buf[:total],
                        log2=np.array([l2w.value, l2h.value, mm.value]), u=u, v=v, colors=colors, mips=mips)

    # Raster vectors: random small triangles in sub-pixel units, with clip rectangles.
    rng = S.Rng(7)
    t = 3000
    corners = rng.integers(t * 6, -2000, 6000).reshape(t, 6).astype(np.int64)
    snap = rng.integers(t * 6, 0, 4).reshape(t, 6)
    corners = np.where(snap == 0, (corners // 256) * 256, corners)  # exact pixel-grid corners
    corners[::17, 2:4] = corners[::17, 0:2]  # coinciding corners
    corners[1::19, 3] = corners[1::19, 1]    # horizontal edges
    corners[2::23, 4] = corners[2::23, 2]    # vertical edges
    clips = np.stack([rng.integers(t, -4, 8), rng.integers(t, -4, 8), rng.integers(t, 1, 24), rng.integers(t, 1, 24)], 1)
    rows = np.zeros((t, 24, 2), dtype=np.int32)
    for i in range(t):
        REF.ref_rasterize(corners[i].ctypes.data_as(C.c_void_p), int(clips[i, 0]), int(clips[i, 1]), int(clips[i, 2]),
                          int(clips[i, 3]), rows[i].ctypes.data_as(C.c_void_p))
    np.savez_compressed(os.path.join(HERE, "raster_vectors.npz"), corners=corners, clips=clips.astype(np.int32), rows=rows)

    # Reciprocal lanes (vendor tagged).
    vals = np.concatenate([np.float32(2.0) ** np.arange(-20, 21, dtype=np.float32), S.Rng(5).uniform(4000, 0.01, 50.0)]).astype(np.float32)
    vals = vals[: (len(vals) // 4) * 4]
    out = np.zeros_like(vals)
    REF.ref_reciprocal(len(vals), vals.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    np.savez_compressed(os.path.join(HERE, "reciprocal.npz"), x=vals, r=out, vendor=np.array(vendor))
    print("done; vendor", vendor)


if __name__ == "__main__":
    main()
