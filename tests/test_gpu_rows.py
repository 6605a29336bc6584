"""The quotient behind every row interval, on the gfx950 library through its C ABI.

A row's side from one edge is clamp((int32)(trunc(n / d) + 1), L, R) (ITriangle2D.cpp:86-148
with the row's clip columns).  The device takes a float estimate and one exact remainder step
for operands below 2^46 and clip columns within +-65536, and the exact division otherwise;
this checks both against Python's integers on quotients at and next to whole numbers, at the
edges of the clamp shortcuts, on large and int32-wrapping quotients and across the 2^46 and
65536 limits.
"""
import numpy as np
import pytest

try:  # import torch at collection time: the first import on a fresh machine is slow
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

pytestmark = pytest.mark.gpu


def exact_side(n, d, L, R):
    q = abs(n) // abs(d)
    if (n < 0) != (d < 0):
        q = -q
    v = q + 1
    v = ((v + (1 << 31)) % (1 << 32)) - (1 << 31)  # the int32 conversion
    return min(max(L, v), R)


def cases(seed=5, count=60000):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = i % 6
        L = int(rng.integers(-1000, 3000))
        R = L + int(rng.integers(-8, 5000))
        if kind == 5:  # clip columns past +-65536
            L, R = int(rng.integers(-80000, -60000)), int(rng.integers(60000, 80000))
        mag = [1 << 8, 1 << 20, 1 << 30, 1 << 40, 1 << 45][int(rng.integers(0, 5))]
        d = int(rng.integers(1, mag + 1)) * (1 if rng.random() < 0.5 else -1)
        if kind in (0, 5):  # quotients inside or next to the clip columns, at or beside whole numbers
            q = int(rng.integers(L - 12, R + 12))
        elif kind == 1:  # at the clamp shortcuts' thresholds
            q = int(rng.choice([R + 2, R + 3, R + 4, R + 5, L - 4, L - 5, L - 6, L - 7]))
        elif kind == 2:  # large quotients, some past int32
            q = int(rng.integers(1 << 20, 1 << 34)) * (1 if rng.random() < 0.5 else -1)
        elif kind == 3:  # around 2^30, where the estimate path hands over to the exact division
            q = int(rng.integers((1 << 30) - 4096, (1 << 30) + 4096)) * (1 if rng.random() < 0.5 else -1)
        else:  # dividends near 2^46
            q = ((1 << 46) + int(rng.integers(-1 << 20, 1 << 20))) // abs(d) * (1 if rng.random() < 0.5 else -1)
        r = int(rng.choice([0, 1, -1, abs(d) - 1, -(abs(d) - 1), int(rng.integers(-abs(d) + 1, abs(d)))]))
        n = q * d + r
        if abs(n) >= (1 << 62):
            continue
        out.append((n, d, L, R))
    return out


def test_edge_side_matches_exact_integers():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from dfpsr_amd import _lib
    L_ = _lib.lib()
    cs = cases()
    n = np.array([c[0] for c in cs], np.int64)
    d = np.array([c[1] for c in cs], np.int64)
    lo = np.array([c[2] for c in cs], np.int32)
    hi = np.array([c[3] for c in cs], np.int32)
    got = np.zeros(len(cs), np.int32)
    print(f"edge side probe: {len(cs)} cases", flush=True)
    rc = L_.dfpsr_hip_probe_edge_side(n.ctypes.data, d.ctypes.data, lo.ctypes.data, hi.ctypes.data, len(cs), got.ctypes.data)
    assert rc == 0
    want = np.array([exact_side(*c) for c in cs], np.int32)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"{bad.size} sides differ, first: " + ", ".join(
        f"n={cs[i][0]} d={cs[i][1]} L={cs[i][2]} R={cs[i][3]} got={got[i]} want={want[i]}" for i in bad[:4])
