"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the checker.

Three kinds of evidence:
  * the reference's own known-answer tests (test/tests/TextureTest.cpp, SimdTest.cpp),
    extracted as data into tests/golden/known_answers.json;
  * outputs of the reference renderer compiled from its sources (tests/golden/*.npz);
  * what the mathematics fixes: closed forms of bilinear weights and box filtering,
    the per-pixel edge-function test, watertight shared edges, the exact depth plane.
Runs on the CPU only.
"""
import json
import os

import numpy as np
import pytest

from dfpsr_amd import scenes as S
from oracle import harness as H
import fixture_scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def vendor():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("vendor_id"):
                return line.split(":")[1].strip()
    except OSError:
        pass
    return "unknown"


@pytest.fixture(scope="module")
def lib():
    return H.load_oracle()


@pytest.fixture(scope="module")
def ka():
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        return json.load(f)


def test_texture_known_answers(lib, ka):
    """TextureTest.cpp: pyramid layout, pixel offsets with wrapping, readPixel, sample_nearest."""
    n = 0
    for block in ka["textures"]:
        writes = np.array(block["writes"] or [[0, 0, 0, 0]], dtype=np.uint32)
        nw = len(block["writes"])
        for c in block["checks"]:
            got = lib.oracle_texture_query(block["log2w"], block["log2h"], block["max_mip"], nw, H.ptr(writes), c["op"],
                                           c.get("x", 0), c.get("y", 0), c.get("mip", 0), c.get("u", 0.0), c.get("v", 0.0))
            assert got == c["expected"], (block["log2w"], block["log2h"], c)
            n += 1
    assert n > 300


def test_linear_mix_known_answer(lib, ka):
    lm = ka["linear_mix"]
    for a, b, w, e in zip(lm["a"], lm["b"], lm["weight"], lm["expected"]):
        assert lib.oracle_mix_colors(a, b, w) == e


def test_linear_mix_closed_form(lib):
    """w = 0 gives a, w = 256 gives b, w = 128 gives floor((a + b) / 2) per channel."""
    rng = S.Rng(3)
    a = rng.u64(500).astype(np.uint32)
    b = rng.u64(500).astype(np.uint32)
    for x, y in zip(a, b):
        x, y = int(x), int(y)
        assert lib.oracle_mix_colors(x, y, 0) == x
        assert lib.oracle_mix_colors(x, y, 256) == y
        half = sum((((x >> s) & 255) + ((y >> s) & 255)) // 2 << s for s in (0, 8, 16, 24))
        assert lib.oracle_mix_colors(x, y, 128) == half
        w = int(rng.integers(1, 0, 257)[0])
        general = sum(((((x >> s) & 255) * (256 - w) + ((y >> s) & 255) * w) >> 8) << s for s in (0, 8, 16, 24))
        assert lib.oracle_mix_colors(x, y, w) == general


def test_reciprocal_known_answer(lib, ka):
    """SimdTest.cpp: reciprocal lanes, compared as that test compares floats (|a - b| < 0.0001,
    noSimd.h:73); the refined rcpps estimate is not exact even on powers of two."""
    x = np.array(ka["reciprocal"]["in"], dtype=np.float32)
    out = np.zeros_like(x)
    lib.oracle_reciprocal(len(x), H.ptr(x), H.ptr(out))
    assert np.all(np.abs(out - np.array(ka["reciprocal"]["expected"], dtype=np.float32)) < 0.0001)


def test_reciprocal_golden(lib):
    g = np.load(os.path.join(GOLDEN, "reciprocal.npz"))
    x = g["x"]
    out = np.zeros_like(x)
    lib.oracle_reciprocal(len(x), H.ptr(x), H.ptr(out))
    # Within 2 ulp of 1/x on any x86 vendor; bit-exact on the vendor that made the golden.
    exact = (1.0 / x.astype(np.float64)).astype(np.float32)
    assert np.all(np.abs(out.view(np.int32) - exact.view(np.int32)) <= 2)
    if str(g["vendor"]) == vendor():
        assert np.array_equal(out, g["r"])


def test_pack_saturated(lib, ka):
    """clampUpper known answer (SimdTest.cpp) and the byte packing of PackOrder.h:191-212."""
    cu = ka["clamp_upper"]
    assert cu["limit"] == pytest.approx(1.5)
    assert lib.oracle_pack_saturated(0.0, 1.9, 254.99, 255.0) == (0 | (1 << 8) | (254 << 16) | (255 << 24))
    assert lib.oracle_pack_saturated(300.0, 255.1, 1e30, 0.4) == (255 | (255 << 8) | (255 << 16) | (0 << 24))
    # NaN clamps to the limit; slightly negative truncates to 0; -1 leaks into every channel.
    assert lib.oracle_pack_saturated(float("nan"), 0.0, 0.0, 0.0) == 255
    assert lib.oracle_pack_saturated(-0.9, 0.0, 0.0, 0.0) == 0
    assert lib.oracle_pack_saturated(-1.0, 0.0, 0.0, 0.0) == 0xFFFFFFFF


def test_pyramid_golden_and_box_filter(lib):
    g = np.load(os.path.join(GOLDEN, "texture_vectors.npz"))
    lv0 = g["level0"]
    cap = 64 * 64 * 2
    buf = np.zeros(cap, dtype=np.uint32)
    l2 = [np.zeros(1, np.int32) for _ in range(3)]
    total = lib.oracle_texture_pyramid(64, 64, 7, H.ptr(lv0), H.ptr(buf), cap, H.ptr(l2[0]), H.ptr(l2[1]), H.ptr(l2[2]))
    assert np.array_equal(buf[:total], g["pyramid"])
    assert [int(a[0]) for a in l2] == list(g["log2"])
    # Closed form: levels smallest first; each level is the truncated mean of 2x2 blocks.
    levels = []
    off = 0
    for m in range(6, -1, -1):
        s = 1 << (6 - m)
        levels.append(buf[off:off + s * s].reshape(s, s))
        off += s * s
    levels = levels[::-1]
    assert np.array_equal(levels[0], lv0)
    for m in range(1, 7):
        src = levels[m - 1].astype(np.uint64)
        out = np.zeros_like(levels[m], dtype=np.uint64)
        for sh in (0, 8, 16, 24):
            ch = (src >> np.uint64(sh)) & np.uint64(255)
            mean = (ch[0::2, 0::2] + ch[1::2, 0::2] + ch[0::2, 1::2] + ch[1::2, 1::2]) // np.uint64(4)
            out |= mean << np.uint64(sh)
        assert np.array_equal(out.astype(np.uint32), levels[m])


def test_bilinear_quads_golden(lib):
    g = np.load(os.path.join(GOLDEN, "texture_vectors.npz"))
    u, v = g["u"], g["v"]
    q = u.shape[0]
    colors = np.zeros((q, 4), np.uint32)
    mips = np.zeros(q, np.uint32)
    lib.oracle_sample_quads(64, 64, 7, H.ptr(g["level0"]), q, H.ptr(u), H.ptr(v), H.ptr(colors), H.ptr(mips))
    assert np.array_equal(mips, g["mips"])
    assert np.array_equal(colors, g["colors"])
    assert len(np.unique(mips)) == 5  # levels 0..4 all exercised


def test_raster_golden(lib):
    g = np.load(os.path.join(GOLDEN, "raster_vectors.npz"))
    for c, clip, rows in zip(g["corners"], g["clips"], g["rows"]):
        got = H.oracle_rasterize(c, *[int(v) for v in clip])
        assert np.array_equal(got, rows[: clip[3]])


def edge_mask(c, L, T, W, Hh):
    """Per-pixel-centre edge test: inside iff value <= threshold on all three edges,
    with threshold -1 on edges that run right-to-left or bottom-to-top (top-left rule)."""
    inside = np.ones((Hh, W), bool)
    ys = (T + np.arange(Hh))[:, None] * 256 + 128
    xs = (L + np.arange(W))[None, :] * 256 + 128
    for k in range(3):
        s, e = c[k], c[(k + 1) % 3]
        thr = -1 if (s[0] > e[0] or (s[0] == e[0] and s[1] > e[1])) else 0
        inside &= ((xs - s[0]) * (e[1] - s[1]) + (ys - s[1]) * (s[0] - e[0])) <= thr
    return inside


def rows_mask(rows, L, W):
    m = np.zeros((rows.shape[0], W), bool)
    for r, (l, rr) in enumerate(rows):
        if rr > l:
            m[r, max(l - L, 0):max(rr - L, 0)] = True
    return m


def test_raster_matches_edge_functions():
    """For pixel columns x >= 1 the row intervals equal the per-pixel edge test.  (Column 0
    differs: the reference divides with truncation toward zero, which the goldens pin.)"""
    rng = S.Rng(11)
    for i in range(600):
        c = rng.integers(6, 0, 24 * 256).reshape(3, 2)
        if len({tuple(p) for p in c}) < 3:
            continue
        L, T, W, Hh = 1, 0, 23, 24
        rows = H.oracle_rasterize(c.reshape(-1), L, T, W, Hh)
        assert np.array_equal(rows_mask(rows, L, W), edge_mask(c, L, T, W, Hh)), c


def test_watertight_shared_edge():
    """Two triangles sharing an edge in opposite directions cover each pixel at most once,
    and their union equals the coverage of the quad they form."""
    rng = S.Rng(12)
    for i in range(400):
        a, b, c, d = rng.integers(8, 256, 20 * 256).reshape(4, 2)
        t1 = np.array([a, b, c])
        t2 = np.array([a, c, d])
        L, T, W, Hh = 1, 0, 22, 22
        m1 = rows_mask(H.oracle_rasterize(t1.reshape(-1), L, T, W, Hh), L, W)
        m2 = rows_mask(H.oracle_rasterize(t2.reshape(-1), L, T, W, Hh), L, W)
        assert not np.any(m1 & m2)


@pytest.fixture(scope="module")
def scene_goldens():
    return np.load(os.path.join(GOLDEN, "scenes.npz"))


def compare_color(got, want, same_vendor):
    if same_vendor:
        return np.array_equal(got, want)
    a = got.view(np.uint8).reshape(-1, 4).astype(int)
    b = want.view(np.uint8).reshape(-1, 4).astype(int)
    ok = np.all(np.abs(a - b) <= 1, axis=1)
    return ok.mean() >= 0.999


@pytest.mark.parametrize("scene", fixture_scenes.fixture_scenes(), ids=lambda s: s.name)
def test_scene_golden(scene, scene_goldens):
    if scene.occluders:
        color, depth, visible = H.oracle_render_queries(scene)
        assert visible.tolist() == scene_goldens[scene.name + "/visible"].tolist()
    else:
        color, depth = H.oracle_render(scene, threads=1)
    want_c = scene_goldens[scene.name + "/color"]
    want_d = scene_goldens[scene.name + "/depth"]
    assert np.array_equal(depth.view(np.uint32), want_d.view(np.uint32))
    same = str(scene_goldens["vendor"]) == vendor() or scene.name.endswith("_id")
    assert compare_color(color, want_c, same)


def test_clip_scenes_exercise_clipping():
    """The clip fixtures really produce sub-triangles: more commands than source triangles
    for some triangle, and corners far outside the canvas are brought inside the clip frustum."""
    for sc in fixture_scenes.clip_scenes():
        tri, flat, bound = H.oracle_commands(sc)
        counts = np.bincount(tri)
        if sc.name != "terrain_small":
            assert counts.max() > 1, sc.name
        # Clip frustum is the view frustum with doubled slopes: |x - W/2| <= W, |y - H/2| <= H.
        x = flat[:, :, 0] / 256.0
        y = flat[:, :, 1] / 256.0
        assert np.all(np.abs(x - sc.width / 2) <= sc.width * 1.01 + 2)
        assert np.all(np.abs(y - sc.height / 2) <= sc.height * 1.01 + 2)


def test_config0_golden():
    g = np.load(os.path.join(GOLDEN, "config0.npz"))
    sc = S.cube_soup()
    color, depth = H.oracle_render(sc, threads=4)
    assert np.array_equal(depth.view(np.uint32), g["depth"].view(np.uint32))
    assert compare_color(color, g["color"], str(g["vendor"]) == vendor())
    cid, _ = H.oracle_render(fixture_scenes.id_color(sc), threads=4)
    assert np.array_equal(cid, g["id_color"])


def test_config1_golden():
    import hashlib
    with open(os.path.join(GOLDEN, "config1.json")) as f:
        meta = json.load(f)
    sc = S.terrain()
    color, depth = H.oracle_render(sc, threads=4)
    assert hashlib.sha256(depth.tobytes()).hexdigest() == meta["depth_sha256"]
    crops = np.load(os.path.join(GOLDEN, "config1_crops.npz"))
    for i in range(5):
        y, x = crops[f"crop{i}_origin"]
        assert np.array_equal(depth[y:y + 64, x:x + 64].view(np.uint32), crops[f"crop{i}_depth"].view(np.uint32))
        if meta["vendor"] == vendor():
            assert np.array_equal(color[y:y + 64, x:x + 64], crops[f"crop{i}_color"])
    if meta["vendor"] == vendor():
        assert hashlib.sha256(color.tobytes()).hexdigest() == meta["color_sha256"]
    cid, did = H.oracle_render(fixture_scenes.id_color(sc), threads=4)
    assert hashlib.sha256(cid.tobytes()).hexdigest() == meta["id_color_sha256"]


def test_threads_identical():
    sc = fixture_scenes.clip_scenes()[2]
    a = H.oracle_render(sc, threads=1)
    b = H.oracle_render(sc, threads=8)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_projection_closed_form():
    """cameraToScreen against the pinhole formula in float64, and flat = trunc(is * 256)."""
    sc = S.terrain(width=320, height=180, grid=32, texture_size=16, resolutions=5)
    pts = sc.points
    cs, is_, flat = H.oracle_project(sc, S.identity_transform(), pts)
    loc = sc.cam_location.astype(np.float64)
    R = loc[3:].reshape(3, 3)
    d = pts.astype(np.float64) - loc[:3]
    cs64 = d @ R.T
    assert np.allclose(cs, cs64, atol=1e-4)
    ws = np.float64(sc.width_slope)
    hs = ws * sc.height / sc.width
    vis = cs64[:, 2] > 0.1
    x = (cs64[:, 0] * (0.5 / ws) / cs64[:, 2] + 0.5) * sc.width
    y = (-cs64[:, 1] * (0.5 / hs) / cs64[:, 2] + 0.5) * sc.height
    assert np.allclose(is_[vis, 0], x[vis], rtol=1e-5, atol=1e-2)
    assert np.allclose(is_[vis, 1], y[vis], rtol=1e-5, atol=1e-2)
    assert np.array_equal(flat[vis], np.trunc(is_[vis].astype(np.float32) * np.float32(256)).astype(np.int64))


def test_depth_is_the_inverse_depth_plane():
    """Stored depth is 1/W interpolated in screen space: within 1e-4 relative of the exact plane
    through the three corners' 1/z (the incremental float chain only adds rounding)."""
    sc = S.single_triangles(64, 48, [[(3.5, 2.5), (60.5, 10.5), (10.5, 44.5)]], depth=[3.0])
    sc.points[1, 2] = 5.0  # tilt the triangle
    sc.points[1, :2] *= 5.0 / 3.0
    color, depth = H.oracle_render(sc)
    cs, is_, flat = H.oracle_project(sc, S.identity_transform(), sc.points)
    A = np.c_[is_.astype(np.float64), np.ones(3)]
    plane = np.linalg.solve(A, 1.0 / cs[:, 2].astype(np.float64))
    ys, xs = np.nonzero(depth > 0)
    exact = plane[0] * (xs + 0.5) + plane[1] * (ys + 0.5) + plane[2]
    assert len(xs) > 500
    assert np.max(np.abs(depth[ys, xs] - exact) / exact) < 1e-4


@pytest.fixture(scope="module")
def depth_goldens():
    return np.load(os.path.join(GOLDEN, "depth_scenes.npz"))


@pytest.mark.parametrize("scene", fixture_scenes.depth_scenes(), ids=lambda s: s.name)
def test_depth_only_golden(scene, depth_goldens):
    """model_renderDepth: the oracle leaves the reference's depth canvas bit for bit."""
    got = H.oracle_render_depth(scene)
    want = depth_goldens[scene.name]
    bad = got.view(np.uint32) != want.view(np.uint32)
    assert not bad.any(), f"{int(bad.sum())} depth pixels differ, first at {np.argwhere(bad)[:3].tolist()}"


def test_depth_only_covers_what_the_renderer_covers():
    """Both paths cut the same row intervals (ITriangle2D::getShape): on a perspective scene
    cleared to zero, the depth-only path writes exactly the pixels the renderer's depth test
    writes, with the same inverse depth up to the order of the plane's additions."""
    checked = 0
    for sc in fixture_scenes.depth_scenes():
        if not sc.perspective or sc.initial_depth is not None:
            continue
        only = H.oracle_render_depth(sc)
        _, full = H.oracle_render(sc, threads=1)
        assert np.array_equal(only > 0, full > 0), sc.name
        np.testing.assert_allclose(only, full, rtol=1e-5, atol=0, err_msg=sc.name)
        checked += 1
    assert checked >= 5


def test_depth_only_keeps_nearest_in_both_projections():
    """Perspective keeps the greater inverse depth, orthogonal the smaller depth: the same
    screen triangle at depth 2 then at depth 4 leaves the depth-2 values wherever it drew."""
    tri = [(2, 2), (30, 2), (2, 30)]
    for perspective in (True, False):
        scenes = [S.single_triangles(32, 32, [tri] * len(d), depth=d) for d in ([2.0], [4.0], [2.0, 4.0])]
        for sc in scenes:
            sc.perspective = perspective
            if not perspective:
                sc.initial_depth = np.full((32, 32), 1.0e6, np.float32)
        near, far, both = (H.oracle_render_depth(sc) for sc in scenes)
        written = near != (0.0 if perspective else 1.0e6)
        assert written.sum() > 50
        assert np.array_equal(both[written], near[written])
        assert np.all((near[written] > far[written]) if perspective else (near[written] < far[written]))


@pytest.mark.reference
@pytest.mark.skipif(not (H.ref_available() and os.path.isdir("/root/reference/Source")),
                    reason="the reference renderer is not built from source here")
@pytest.mark.parametrize("index", [0, 1])
def test_depth_only_special_canvas_values_match_reference(index):
    """model_renderDepth over special canvas values (perspective keeps the greater, orthogonal
    the smaller): the oracle and the compiled reference leave the same bits."""
    sc = fixture_scenes.special_depth_canvas(fixture_scenes.depth_scenes()[index], 100 + index)
    got = H.oracle_render_depth(sc)
    want = H.ref_render_depth(sc)
    assert (got.view(np.uint32) != sc.initial_depth.view(np.uint32)).sum() > 50
    bad = got.view(np.uint32) != want.view(np.uint32)
    assert not bad.any(), f"{int(bad.sum())} depth pixels differ, first at {np.argwhere(bad)[:3].tolist()}"


def test_projected_scenes_move_coverage():
    """The ProjectedPoint fixtures really give corners whose flat and is differ from
    cameraToScreen(cs): their canvases differ from the same triangles projected by the
    reference's own cameraToScreen (the oracle's, pinned by the goldens above)."""
    import copy
    for sc in fixture_scenes.projected_scenes():
        world = sc.points[sc.indices[:, :3]].reshape(-1, 3)
        cs, is_, flat = H.oracle_project(sc, S.identity_transform(), world)
        assert np.allclose(cs.reshape(-1, 3, 3), sc.projected["cs"], atol=1e-5), sc.name
        exact = copy.copy(sc)
        rec = sc.projected.copy()
        rec["is"], rec["flat"] = is_.reshape(-1, 3, 2), flat.reshape(-1, 3, 2)
        exact.projected = rec
        moved = int((rec["flat"] != sc.projected["flat"]).any(-1).sum())
        assert moved > sc.polygon_count // 2, sc.name
        a = H.oracle_render(sc)
        b = H.oracle_render(exact)
        assert not np.array_equal(a[0], b[0]), sc.name


def test_occlusion_scenes_exercise_the_grid():
    """Each occluder fixture drops draw commands (completeOcclusion) or whole models, the box
    wall changes the image, and the box queries answer both ways over the set."""
    import copy
    answers = []
    for sc in fixture_scenes.occlusion_scenes():
        plain = copy.copy(sc)
        plain.occluders = None
        with_occluders = len(H.oracle_commands(sc)[0])
        without = len(H.oracle_commands(plain)[0])
        grid, occluded = H.oracle_occlusion_grid(sc)
        assert occluded and np.isfinite(grid).any(), sc.name
        assert with_occluders <= without, sc.name
        answers += H.oracle_render_queries(sc)[2].tolist()
        if sc.name in ("occ_box", "occ_existing", "occ_top_rows_orthogonal"):
            assert with_occluders < without * 0.8, sc.name
    assert 0 in answers and 1 in answers


def test_occlusion_grid_cells_follow_the_rules():
    """Mathematical checks of the grid the occluders leave: a box occluder's cells hold its
    farthest corner depth or +inf; top-row cells of a constant inverse depth d hold exactly
    1/d; existing triangles never write below the nearest queued corner."""
    import copy
    sc = fixture_scenes.occlusion_scenes()[0]
    grid, _ = H.oracle_occlusion_grid(sc)
    values = set(grid[np.isfinite(grid)].tolist())
    assert len(values) == 1
    flat = copy.copy(fixture_scenes.occlusion_scenes()[1])
    flat.initial_depth = np.full((flat.height, flat.width), 0.25, np.float32)
    flat.occluders = [("top_rows", 0)]
    grid, _ = H.oracle_occlusion_grid(flat)
    assert np.all(grid == np.float32(4.0))
