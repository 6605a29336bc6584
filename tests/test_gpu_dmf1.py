"""DMF1 models on the GPU: the import fetches each part's textures from the caller's pool as
the reference's import does (tests/golden/dmf1.npz records which names the reference asked
for, per part), and an imported sample model draws what the oracle draws from the reference's
own import of it."""
import os

import numpy as np
import pytest

try:  # import torch at collection time: the first import on a fresh machine is slow
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

from dfpsr_amd import scenes as S
from oracle import harness as H

pytestmark = pytest.mark.gpu

GOLDEN = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dmf1.npz"))


@pytest.fixture(scope="module")
def api():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from dfpsr_amd import api as A
    A.lib()
    return A


def golden(name, detail):
    g = {k: GOLDEN[f"{name}/{detail}/{k}"] for k in ("points", "parts", "indices", "colors", "texcoords", "names")}
    lines = bytes(g["names"]).decode().split("\n")[:-1]
    g["texture_names"] = lines[len(g["parts"]):]
    return g


def pool_for(api, names, size=64):
    """Noise textures under the given names: diffuse maps with 5 resolutions and light maps
    with 1, as BasicResourcePool makes them for the import (Model.cpp:244-265)."""
    pool, textures = {}, {}
    for k, n in enumerate(names):
        t = S.noise_checker_texture(size, 900 + k, 5, check=8)
        tex = api.texture_create_RgbaU8(size, size, 5)
        api.texture_write_level0(tex, t.level0)
        api.texture_generatePyramid(tex)
        pool[n] = tex
        textures[n] = t
    return pool, textures


@pytest.mark.parametrize("name", ["Model_Test", "Model_Crate", "Model_Barrel", "random_model"])
def test_import_fetches_the_reference_textures(api, name):
    for detail in (0, 1, 2):
        g = golden(name, detail)
        pool, _ = pool_for(api, g["texture_names"] + ["NotAsked"], size=4)
        # Pool names in another case: the reference's pool compares names ignoring case.
        pool = {(k.upper() if i % 2 else k.lower()): v for i, (k, v) in enumerate(pool.items())}
        handle_of = {v.handle.value if hasattr(v.handle, "value") else v.handle: k.lower() for k, v in pool.items()}
        m = api.model_importDMF1(bytes(GOLDEN[name + "/text"]), detail, pool)
        for p, (_, diffuse, light) in enumerate(g["parts"]):
            part = api.model_getPart(m, p)
            for got, number in ((part["diffuse"], diffuse), (part["light"], light)):
                want = g["texture_names"][number - 1].lower() if number > 0 else None
                assert (handle_of[got] if got else None) == want, (name, detail, p)


def test_imported_crate_draws_as_the_oracle_draws_it(api):
    """Model_Crate imported on the device, textured from a pool, drawn by renderer_giveTask;
    the oracle draws the same scene built from the reference's import."""
    name = "Model_Crate"
    g = golden(name, 2)
    pool, textures = pool_for(api, g["texture_names"])
    m = api.model_importDMF1(bytes(GOLDEN[name + "/text"]), 2, pool)
    n = g["indices"].shape[0]
    indices = np.full((n, 4), -1, np.int32)
    indices[:, :3] = g["indices"]
    colors = np.zeros((n, 4, 4), np.float32)
    colors[:, :3] = g["colors"]
    tex = np.zeros((n, 4, 4), np.float32)
    tex[:, :3] = g["texcoords"]
    parts, first = [], 0
    names = g["texture_names"]
    for count, diffuse, light in g["parts"]:
        parts.append(S.Part(first, first + int(count), diffuse=int(diffuse) - 1, light=int(light) - 1))
        first += int(count)
    sc = S.Scene(name="crate", width=320, height=240, points=g["points"], indices=indices, colors=colors, texcoords=tex,
                 parts=parts, textures=[textures[t] for t in names],
                 cam_location=S.look_transform((0.15, 0.2, -0.35), -0.4, 0.5), width_slope=1.0)
    ref_color, ref_depth = H.oracle_render(sc)
    assert (ref_depth > 0).mean() > 0.3
    r = api.renderer_create()
    color = np.zeros((sc.height, sc.width), np.uint32)
    depth = np.zeros((sc.height, sc.width), np.float32)
    cam = api.camera_create(True, sc.cam_location, sc.width, sc.height, sc.width_slope, sc.near_clip, sc.far_clip)
    api.renderer_begin(r, api.host_image(color), api.host_image(depth))
    api.renderer_giveTask(r, m, S.identity_transform(), cam)
    api.renderer_end(r)
    assert np.array_equal(depth.view(np.uint32), ref_depth.view(np.uint32))
    a = color.view(np.uint8).reshape(-1, 4).astype(int)
    b = ref_color.view(np.uint8).reshape(-1, 4).astype(int)
    bad = np.any(np.abs(a - b) > 1, axis=1)
    assert bad.sum() <= int(0.001 * bad.size)
