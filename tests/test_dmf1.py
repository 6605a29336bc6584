"""The DMF1 model import (dfpsr_hip_model_import_dmf1, C++14 host code in the library) against
what the reference's importFromContent_DMF1 made of the same texts (tests/golden/dmf1.npz,
written by tests/golden/make_dmf1_goldens.py from the compiled reference): the SDK's three
sample models and synthetic texts for the tokenizer, at detail levels 0, 1 and 2.  Points,
polygons, corner colours and texture coordinates must match bit for bit.  These run on the
CPU: importing touches no device memory.  Texture assignment needs device textures and is
checked in tests/test_gpu_dmf1.py."""
import os

import numpy as np
import pytest

from dfpsr_amd import _lib

GOLDEN = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dmf1.npz"))
CASES = [(str(n), d) for n in GOLDEN["names"] for d in (0, 1, 2)]


@pytest.fixture(scope="module")
def api():
    from dfpsr_amd import api as A
    A.lib()
    return A


def golden(name, detail):
    g = {k: GOLDEN[f"{name}/{detail}/{k}"] for k in ("points", "parts", "indices", "colors", "texcoords", "names")}
    lines = bytes(g["names"]).decode().split("\n")[:-1]
    parts = g["parts"]
    g["part_names"] = lines[:len(parts)]
    g["texture_names"] = lines[len(parts):]
    return g


@pytest.mark.parametrize("name,detail", CASES, ids=[f"{n}-{d}" for n, d in CASES])
def test_import_matches_reference(api, name, detail):
    text = bytes(GOLDEN[name + "/text"])
    g = golden(name, detail)
    if name == "no_header":
        with pytest.raises(api.DfpsrError) as e:
            api.model_importDMF1(text, detail)
        assert e.value.code == _lib.ERROR_ARGUMENT
        return
    m = api.model_importDMF1(text, detail)
    assert np.array_equal(api.model_getPoints(m).view(np.uint32), g["points"].view(np.uint32))
    assert api.model_getNumberOfParts(m) == len(g["parts"])
    first = 0
    for p, (count, _, _) in enumerate(g["parts"]):
        part = api.model_getPart(m, p)
        assert api.model_getPartName(m, p) == g["part_names"][p]
        assert part["indices"].shape[0] == count
        sl = slice(first, first + count)
        assert np.array_equal(part["indices"][:, :3], g["indices"][sl])
        assert np.all(part["indices"][:, 3] == -1)
        assert np.array_equal(part["colors"][:, :3].view(np.uint32), g["colors"][sl].view(np.uint32))
        assert np.array_equal(part["texcoords"][:, :3].view(np.uint32), g["texcoords"][sl].view(np.uint32))
        assert part["diffuse"] is None and part["light"] is None  # no pool given
        first += count
    assert first == g["indices"].shape[0]


def test_byte_order_mark_is_skipped(api):
    text = bytes(GOLDEN["Model_Test/text"])
    a = api.model_importDMF1(text, 2)
    b = api.model_importDMF1(b"\xef\xbb\xbf" + text, 2)
    assert np.array_equal(api.model_getPoints(a), api.model_getPoints(b))


def test_sample_models_exercise_detail_levels_and_merging():
    """The barrel's detail levels select different parts, and points are shared between
    triangles (merged within 0.00001)."""
    low, high = golden("Model_Barrel", 0), golden("Model_Barrel", 2)
    assert low["indices"].shape[0] < high["indices"].shape[0]
    for name in ("Model_Test", "Model_Crate", "Model_Barrel"):
        g = golden(name, 2)
        assert g["points"].shape[0] < 3 * g["indices"].shape[0]
