"""Shared test helpers: bitwise comparison, brute-force pins, golden loading."""

import functools
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def canonical(a):
    """Canonicalise -0.0 -> +0.0 and every NaN to one payload for bitwise
    comparison (the transforms' bit-exact bar)."""
    a = np.array(a, copy=True)
    if a.dtype.kind == "f":
        a[a == 0] = 0
        a[np.isnan(a)] = np.nan
    return a


def bitwise_equal(a, b):
    a, b = canonical(a), canonical(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


def assert_bitwise(a, b, what=""):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    assert a.dtype == b.dtype, (what, a.dtype, b.dtype)
    if not bitwise_equal(a, b):
        ca, cb = canonical(a).ravel(), canonical(b).ravel()
        width = ca.dtype.itemsize
        ua = ca.view(np.uint64 if width == 8 else np.uint32)
        ub = cb.view(np.uint64 if width == 8 else np.uint32)
        mism = np.flatnonzero(ua != ub)
        idx = np.unravel_index(mism[0], a.shape)
        raise AssertionError(
            f"{what}: {mism.size} of {a.size} elements differ; first at {idx}: "
            f"{ca[mism[0]]!r} vs {cb[mism[0]]!r}"
        )


def canonical_sha256(a):
    a = np.ascontiguousarray(a)
    a = np.where(a == 0, np.zeros((), a.dtype), a)
    return hashlib.sha256(a.tobytes()).hexdigest()


@functools.lru_cache(maxsize=None)
def golden_small():
    with np.load(os.path.join(GOLDEN, "golden_small.npz")) as f:
        return {k: f[k] for k in f.files}


@functools.lru_cache(maxsize=None)
def golden_hashes():
    with open(os.path.join(GOLDEN, "golden_hashes.json")) as f:
        return json.load(f)


# --------------------------------------------------------------------------
# Brute-force digital lines (Press 2006 / Brady 1998 definition).
# A digital line across m rows that starts at offset h and drops by d: the
# top half drops floor(d/2) from h; the bottom half starts at h + ceil(d/2)
# and also drops floor(d/2).  For an m x m image the four quadrants view the
# image as a, a.T, flipud(a).T and flipud(a); output row k holds the line
# starting at offset h = m - 1 - k, output column the drop d.
# --------------------------------------------------------------------------


def digital_line(m, h, d):
    if m == 1:
        return [(0, h)]
    half = m // 2
    lo = d // 2
    top = digital_line(half, h, lo)
    bottom = digital_line(half, h + (d - lo), lo)
    return top + [(i + half, j) for i, j in bottom]


def brute_force_adrt(img):
    img = np.asarray(img)
    m = img.shape[0]
    views = [img, img.T, np.flipud(img).T, np.flipud(img)]
    out = np.zeros((4, 2 * m - 1, m), dtype=np.float64)
    for q, v in enumerate(views):
        for k in range(2 * m - 1):
            h = m - 1 - k
            for d in range(m):
                tot = 0.0
                for i, j in digital_line(m, h, d):
                    if 0 <= i < m and 0 <= j < m:
                        tot += float(v[i, j])
                out[q, k, d] = tot
    return out


def make_unaligned(a):
    b = np.frombuffer(b"\x00" + a.tobytes("C"), dtype=a.dtype, offset=1).reshape(a.shape)
    assert not b.flags.aligned
    return b
