"""Writes tests/golden/dmf1.npz: DMF1 model texts and what the reference's
importFromContent_DMF1 makes of them at detail levels 0, 1 and 2.

Run in the build container only (it needs /root/reference and oracle/_ref/libdfpsr_ref.so,
built by `make -f oracle/Makefile.ref`):

    python tests/golden/make_dmf1_goldens.py

Inputs (stored as data, UTF-8 bytes): the SDK's three sample models
(Source/SDK/cube/media/Model_{Test,Crate,Barrel}.dmf) and synthetic texts written here that
exercise the tokenizer: indices, '~' and repeated signs, decimal commas, exponents, CR LF and
tabs, case-insensitive names, unknown and misplaced namespaces, out-of-range texture and
vertex indices, over-long names, unfinished statements, points merged within the 0.00001
threshold and kept apart just outside it, and parts limited to detail levels.

Outputs per (text, detail level): points, per part (polygon count, diffuse and light texture
numbers -- 1 + the index of the name among the names the import fetched, 0 for none), each
polygon's three corners (point index, RGBA, u1 v1 u2 v2), part names and fetched names.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import harness  # noqa: E402

REF = harness.load_ref()
MEDIA = "/root/reference/Source/SDK/cube/media"


def synthetic_texts():
    rng = np.random.default_rng(2024)
    out = {}

    def num(v, style):
        if style == 0:
            return f"{v:.6f}"
        if style == 1:
            return f"{v:.3f}".replace(".", ",")
        if style == 2:
            return ("~" if v < 0 else "") + f"{abs(v):.5f}"
        if style == 3:
            return f"{v * 100:.4f}e-2"
        return f"{v:.9g}"

    # A random model over three parts with every property written several ways.
    lines = ["DMF1", "FilterType(Alpha)", "CullingType(Radius)", "BoundMultiplier(1.5)"]
    shaders = ["M_Diffuse_0Tex", "m_diffuse_1tex", "M_DIFFUSE_2TEX", "M_Unknown"]
    base = rng.uniform(-2, 2, size=(40, 3))
    for part in range(4):
        lines.append("<Part>" if part % 2 == 0 else "<pArT>")
        lines.append(f"Name(Part number {part})" if part != 2 else "name(second_part)")
        lines.append(f"Shader[0]({shaders[part]}) Shader[1](M_Shadow_Solid)")
        lines.append(f"Texture[0](Tex{part}A)\tTexture[1](Tex{part}B) Texture[16](Ignored) Texture[1.6](RoundedIndex{part})")
        if part == 1:
            lines.append("MinDetailLevel(1) MaxDetailLevel(2)")
        if part == 3:
            lines.append("MaxDetailLevel(0.4)")
        for t in range(int(rng.integers(6, 14))):
            lines.append("<Triangle>" if t % 3 else "<TRIANGLE>")
            for v in range(3):
                p = base[int(rng.integers(0, len(base)))] + rng.choice([0.0, 3e-6, 2e-5, -4e-6], size=3)
                style = int(rng.integers(0, 5))
                props = [f"X[{v}]({num(p[0], style)})", f"Y[{v}]({num(p[1], style)})", f"z[{v}]({num(p[2], style)})"]
                if rng.random() < 0.7:
                    props += [f"CR[{v}]({num(rng.uniform(0, 1), 0)})", f"CG[{v}]({num(rng.uniform(0, 1), 1)})",
                              f"CB[{v}]({num(rng.uniform(0, 1), 4)})", f"CA[{v}]({num(rng.uniform(0, 1), 0)})"]
                if rng.random() < 0.7:
                    props += [f"U1[{v}]({num(rng.uniform(-1, 2), 3)})", f"V1[{v}]({num(rng.uniform(-1, 2), 0)})",
                              f"U2[{v}]({num(rng.uniform(-1, 2), 2)})", f"V2[{v}]({num(rng.uniform(-1, 2), 4)})"]
                rng.shuffle(props)
                lines.append(" ".join(props))
            if t == 2:
                lines.append("X[3](99) Y[-1](99) Unknown[0](5)")
    lines.append("<Bone> Name(b) <Triangle> X[0](1)")
    lines.append("<Shape> <Point> X(1) Y(2)")
    out["random_model"] = "\r\n".join(lines) + "\r\n"

    # Tokenizer corner cases on one part.
    tricky = [
        "DMF1 <Part> Name(tricky) Shader(M_Diffuse_1Tex) Texture(Brick)",
        "<Triangle> X[0](1.5e2) Y[0](--2) Z[0](3e-1) X[1](-0.5) Y[1](1,25) Z[1](~~7) X[2](4) Y[2](5) Z[2](6)",
        "<Triangle> X[0](1) X[0](2) Y[0](3) [1] X(9) X[2](0.0000049) Y[2](0) Z[2](0)",
        "<Triangle> X[0](1 2) Y[0]( 3 ) X[1](a<b) Z[2](7)",
        "<" + "N" * 300 + "> <Triangle> X[0](8)",
        "<Part> Name(" + "long" * 80 + ") <Triangle> X[1](2)",
        "<Triangle> X[0](0.1) Y[0](0.2) Z[0](0.3) X[1](0.10000001) Y[1](0.2) Z[1](0.3) X[2](0.1) Y[2](0.20002) Z[2](0.3)",
        "X[0](5",
    ]
    out["tricky"] = "\n".join(tricky)
    out["no_header"] = "DMF2 <Part> <Triangle> X[0](1)"
    out["empty_model"] = "DMF1"
    return out


def import_ref(text: bytes, detail: int):
    cap_points, cap_parts, cap_polys, cap_names = 200000, 256, 200000, 1 << 20
    points = np.zeros((cap_points, 3), np.float32)
    point_count = C.c_int32()
    part_info = np.zeros((cap_parts, 3), np.int32)
    part_count = C.c_int32()
    indices = np.zeros((cap_polys, 3), np.int32)
    colors = np.zeros((cap_polys, 3, 4), np.float32)
    tex = np.zeros((cap_polys, 3, 4), np.float32)
    names = C.create_string_buffer(cap_names)
    REF.ref_import_dmf1.restype = C.c_int32
    n = REF.ref_import_dmf1(C.c_char_p(text), detail, cap_points, points.ctypes.data_as(C.c_void_p), C.byref(point_count),
                            cap_parts, part_info.ctypes.data_as(C.c_void_p), C.byref(part_count), cap_polys,
                            indices.ctypes.data_as(C.c_void_p), colors.ctypes.data_as(C.c_void_p), tex.ctypes.data_as(C.c_void_p),
                            cap_names, names)
    assert n >= 0, n
    return {"points": points[:point_count.value].copy(), "parts": part_info[:part_count.value].copy(),
            "indices": indices[:n].copy(), "colors": colors[:n].copy(), "texcoords": tex[:n].copy(),
            "names": np.frombuffer(names.value, np.uint8).copy()}


def main():
    texts = {}
    for name in ("Model_Test", "Model_Crate", "Model_Barrel"):
        with open(os.path.join(MEDIA, name + ".dmf"), "rb") as f:
            raw = f.read()
        if raw.startswith(b"\xef\xbb\xbf"):
            raw = raw[3:]  # string_load decodes the byte order mark away
        texts[name] = raw
    for name, text in synthetic_texts().items():
        texts[name] = text.encode("ascii")
    data = {"names": np.array(list(texts.keys()))}
    for name, raw in texts.items():
        data[name + "/text"] = np.frombuffer(raw, np.uint8)
        for detail in (0, 1, 2):
            for key, value in import_ref(raw, detail).items():
                data[f"{name}/{detail}/{key}"] = value
        print(name, len(raw), "bytes;", [len(data[f"{name}/{d}/indices"]) for d in (0, 1, 2)], "polygons")
    np.savez_compressed(os.path.join(HERE, "dmf1.npz"), **data)


if __name__ == "__main__":
    main()
