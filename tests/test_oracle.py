"""Pins for the CPU oracle (runs without a GPU).

Each oracle function is checked against something other than its own
formula: the compiled reference (bit for bit, when oracle/_ref was built),
the committed golden vectors, brute-force digital-line sums, materialised
adjointness, closed forms and fixed points.  A planted-mistake test shows
that a plausible error (floor instead of ceil in the stage shift) is caught.
"""

import itertools

import numpy as np
import pytest

from helpers import (
    assert_bitwise,
    brute_force_adrt,
    canonical_sha256,
    golden_hashes,
    golden_small,
)
from oracle import oracle as O
from oracle import ref as R

DTYPES = ["float32", "float64"]
SIZES = [1, 2, 4, 8, 16]


# ---- golden vectors ---------------------------------------------------------


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("n", SIZES)
def test_golden_small(dt, n):
    g = golden_small()
    key = f"{dt}_n{n}"
    img, dat = g[f"{key}_img"], g[f"{key}_dat"]
    assert_bitwise(O.adrt(img), g[f"{key}_adrt"], "adrt")
    assert_bitwise(O.bdrt(dat), g[f"{key}_bdrt"], "bdrt")
    assert_bitwise(O.iadrt(dat), g[f"{key}_iadrt"], "iadrt")
    assert_bitwise(O.iadrt(O.adrt(np.round(img * 8))), g[f"{key}_iadrt_of_adrt"], "iadrt(adrt)")
    for s in range(O.num_iters(n)):
        assert_bitwise(O.adrt_step(dat, s), g[f"{key}_adrt_step{s}"], f"adrt_step {s}")
        assert_bitwise(O.bdrt_step(dat, s), g[f"{key}_bdrt_step{s}"], f"bdrt_step {s}")
    if n >= 2:
        assert_bitwise(O.fmg_restriction(dat), g[f"{key}_restriction"], "restriction")
    hp = g[f"{key}_hp_in"]
    assert_bitwise(O.fmg_highpass(hp), g[f"{key}_highpass"], "highpass")
    assert_bitwise(O.fmg_prolongation(hp), g[f"{key}_prolongation"], "prolongation")


def test_golden_hashes():
    h = golden_hashes()
    x = np.random.default_rng(0).random((8, 128, 128), dtype=np.float32)
    y = O.adrt(x)
    assert canonical_sha256(y) == h["adrt_f32_8x128_seed0"]
    assert canonical_sha256(O.bdrt(y)) == h["bdrt_f32_of_adrt_8x128_seed0"]
    assert canonical_sha256(O.iadrt(y)) == h["iadrt_f32_of_adrt_8x128_seed0"]
    x64 = np.random.default_rng(0).random((2, 256, 256))
    assert canonical_sha256(O.adrt(x64)) == h["adrt_f64_2x256_seed0"]
    big = np.random.default_rng(1).random((1, 1024, 1024), dtype=np.float32)
    assert canonical_sha256(O.adrt(big)) == h["adrt_f32_1x1024_seed1"]


# ---- compiled reference -------------------------------------------------------

needs_ref = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built")


@needs_ref
@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("n", [1, 2, 32, 64, 256])
def test_matches_compiled_reference(dt, n):
    rng = np.random.default_rng(n)
    img = (rng.standard_normal((3, n, n)) * 10).astype(dt)
    img[0, 0, 0] = -0.0
    dat = rng.standard_normal((2, 4, 2 * n - 1, n)).astype(dt)
    raw = lambda a: np.ascontiguousarray(a).view(np.uint8)  # noqa: E731
    # Bit-exact including signed zeros.
    assert np.array_equal(raw(O.adrt(img)), raw(R.adrt(img)))
    assert np.array_equal(raw(O.bdrt(dat)), raw(R.bdrt(dat)))
    assert np.array_equal(raw(O.iadrt(dat)), raw(R.iadrt(dat)))
    for s in range(O.num_iters(n)):
        assert np.array_equal(raw(O.adrt_step(dat, s)), raw(R.adrt_step(dat, s)))
        assert np.array_equal(raw(O.bdrt_step(dat, s)), raw(R.bdrt_step(dat, s)))
    if n >= 2:
        assert np.array_equal(raw(O.fmg_restriction(dat)), raw(R.fmg_restriction(dat)))
        hp = rng.standard_normal((2, n, n + 3)).astype(dt)
        assert np.array_equal(raw(O.fmg_highpass(hp)), raw(R.fmg_highpass(hp)))
        assert np.array_equal(raw(O.fmg_prolongation(hp)), raw(R.fmg_prolongation(hp)))


# ---- brute force / mathematics ----------------------------------------------


@pytest.mark.parametrize("n", SIZES)
def test_adrt_brute_force_integer(n):
    rng = np.random.default_rng(100 + n)
    img = rng.integers(-50, 50, size=(n, n)).astype(np.float64)
    assert np.array_equal(O.adrt(img), brute_force_adrt(img))


def _numpy_adrt(img, shift):
    """Stage chain in NumPy with a pluggable shift rule (for the planted-mistake test)."""
    n = img.shape[-1]
    buf = np.zeros((4, n, 1, 2 * n - 1))
    # Quadrant loads of adrt_cdefs_adrt.hpp:123-175 (row r, offset c).
    buf[:, :, 0, :n] = np.stack([img[:, ::-1], img[::-1, :].T, img.T, img[::-1, ::-1]])
    while buf.shape[1] > 1:
        R_, S_ = buf.shape[1] // 2, buf.shape[2] * 2
        out = np.zeros((4, R_, S_, 2 * n - 1))
        for r in range(R_):
            for a in range(S_):
                sh = shift(a)
                x, y = buf[:, 2 * r, a // 2], buf[:, 2 * r + 1, a // 2]
                out[:, r, a, :] = x
                out[:, r, a, sh:] += y[:, : 2 * n - 1 - sh]
        buf = out
    return buf[:, 0].transpose(0, 2, 1)


def test_planted_shift_mistake_is_caught():
    n = 8
    img = np.random.default_rng(5).integers(0, 9, size=(n, n)).astype(np.float64)
    good = _numpy_adrt(img, lambda a: (a + 1) // 2)
    bad = _numpy_adrt(img, lambda a: a // 2)
    bf = brute_force_adrt(img)
    assert np.array_equal(good, bf)
    assert np.array_equal(O.adrt(img), good)
    assert not np.array_equal(bad, bf)


def test_planted_quadrant_mistake_is_caught():
    img = np.random.default_rng(6).integers(0, 9, size=(8, 8)).astype(np.float64)
    bf = brute_force_adrt(img)
    swapped = O.adrt(img.T)  # transposed operand
    assert np.array_equal(O.adrt(img), bf)
    assert not np.array_equal(swapped, bf)


@pytest.mark.parametrize("n", [1, 2, 4, 8])
def test_materialized_adjoint(n):
    full = n * n
    size = 4 * n * (2 * n - 1)
    a_in = np.eye(full).reshape(full, n, n)
    b_in = np.eye(size).reshape(size, 4, 2 * n - 1, n)
    a_mat = O.adrt(a_in).reshape(full, -1).T
    b_mat = O.adjoint(b_in).reshape(size, full)
    assert np.array_equal(a_mat, b_mat)


@pytest.mark.parametrize("n", [2, 4, 8])
def test_step_transposes(n):
    size = 4 * n * (2 * n - 1)
    basis = np.eye(size).reshape(size, 4, 2 * n - 1, n)
    L = O.num_iters(n)
    for k in range(L):
        b = O.bdrt_step(basis, k).reshape(size, size)
        a = O.adrt_step(basis, L - 1 - k).reshape(size, size)
        assert np.array_equal(a, b.T)


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32])
def test_iadrt_round_trip_integer(n):
    x = np.random.default_rng(n).integers(-9, 9, size=(2, n, n)).astype(np.float64)
    inv = O.truncate(O.iadrt(O.adrt(x))).mean(axis=-3)
    assert np.array_equal(inv, x)


def test_adrt_equals_step_chain():
    x = np.random.default_rng(3).standard_normal((2, 32, 32)).astype(np.float32)
    a = O.adrt_init(x)
    for s in range(O.num_iters(32)):
        a = O.adrt_step(a, s)
    assert_bitwise(a, O.adrt(x))


def test_bdrt_equals_step_chain_values():
    y = np.random.default_rng(4).standard_normal((4, 63, 32))
    a = y
    for s in range(O.num_iters(32)):
        a = O.bdrt_step(a, s)
    assert np.allclose(a, O.bdrt(y), rtol=0, atol=1e-12)


def test_spot_checks_vertical_line():
    x = np.zeros((8, 8))
    x[:, 0] = 1
    y = O.adrt(x)
    assert y[0, 7, 0] == 8 and np.count_nonzero(y[0, :, 0]) == 1
    assert np.all(y[0, -8:, -1] == 1) and np.all(y[0, :7, -1] == 0)
    assert np.all(y[2, :8, 0] == 1) and np.all(y[2, -7:, 0] == 0)
    assert y[3, 7, 0] == 8 and np.count_nonzero(y[3, :, 0]) == 1


def test_init_truncate_inverse():
    for dt in ("float32", "float64", "int32", "int64"):
        x = np.arange(3 * 16 * 16).reshape(3, 16, 16).astype(dt)
        assert np.array_equal(O.truncate(O.adrt_init(x)), np.repeat(x[:, None], 4, axis=1))


def test_fmg_closed_forms():
    rng = np.random.default_rng(9)
    a = rng.standard_normal((2, 4, 11, 6))
    assert np.allclose(O.fmg_restriction(a), 0.25 * (a[..., 0:-1:2, ::2] + a[..., 1::2, ::2]))
    b = rng.standard_normal((3, 5, 6))
    assert np.array_equal(O.fmg_prolongation(b), np.repeat(np.repeat(b, 2, -1), 2, -2))
    k = np.array([[-1 / 16, -1 / 8, -1 / 16], [-1 / 8, 3 / 4, -1 / 8], [-1 / 16, -1 / 8, -1 / 16]])
    c = rng.standard_normal((7, 9))
    win = np.lib.stride_tricks.sliding_window_view(np.pad(c, 1, mode="reflect"), (3, 3))
    assert np.allclose(O.fmg_highpass(c), (win * k).sum(axis=(-1, -2)))


@pytest.mark.parametrize("dt", DTYPES)
def test_highpass_fixed_points(dt):
    x, y = np.meshgrid(np.arange(7), np.arange(10))
    cb = ((x % 2) ^ (y % 2)).astype(dt) * 2 - 1
    assert np.array_equal(O.fmg_highpass(cb), cb)
    st = (y % 2).astype(dt) * 2 - 1
    assert np.array_equal(O.fmg_highpass(st), st)
    assert np.array_equal(O.fmg_highpass(np.full((14, 5), 5, dtype=dt)), np.zeros((14, 5), dt))


def _press_inverse(arr):
    """Recursive Press FMG (Press 2006): restrict, solve, prolong, correct."""
    if arr.shape[-1] == 1:
        return arr[0]
    coarse = 0.25 * (arr[:, 0:-1:2, ::2] + arr[:, 1::2, ::2])
    f = np.repeat(np.repeat(_press_inverse(coarse), 2, -1), 2, -2)
    corr = O.truncate(O.bdrt(O.adrt(f) - arr)).mean(axis=0) / (arr.shape[-1] - 1)
    k = np.array([[-1 / 16, -1 / 8, -1 / 16], [-1 / 8, 3 / 4, -1 / 8], [-1 / 16, -1 / 8, -1 / 16]])
    win = np.lib.stride_tricks.sliding_window_view(np.pad(corr, 1, mode="reflect"), (3, 3))
    return f - (win * k).sum(axis=(-1, -2))


@pytest.mark.parametrize("dt", DTYPES)
def test_fmg_step_vs_recursive(dt):
    arr = np.arange(4 * 31 * 16).reshape(4, 31, 16).astype(dt)
    assert np.allclose(O.iadrt_fmg_step(arr), _press_inverse(arr.astype(np.float64)), atol=1e-3)


@pytest.mark.parametrize("dt", DTYPES)
def test_iadrt_fmg_round_trip(dt):
    orig = np.arange(64).reshape(8, 8).astype(dt)
    inv = O.iadrt_fmg(O.adrt(orig), max_iters=50)
    assert inv.dtype == orig.dtype and np.allclose(inv, orig, atol=1e-3)


def test_fmg_iter_monotone():
    steps = np.linspace(-10, 10, 32)
    x, y = np.meshgrid(steps, steps)
    img = np.exp(-np.hypot(x, y))
    it = itertools.islice(O.iadrt_fmg_iter(O.adrt(img)), 10)
    res = [np.linalg.norm(img - v) for v in it]
    assert np.all(np.diff(res) <= 0)


def test_normal_operator_symmetric_psd():
    rng = np.random.default_rng(11)
    u, v = rng.standard_normal((2, 16, 16))
    mu, mv = O.normal_op(u), O.normal_op(v)
    assert np.isclose(np.vdot(mu, v), np.vdot(u, mv), rtol=1e-12)
    assert np.vdot(mu, u) > 0


def test_cg_normal_converges():
    img = np.random.default_rng(2).random((16, 16))
    x, k, hist = O.cg_normal(O.adrt(img), rtol=1e-8, maxiter=500)
    assert hist[-1] <= 1e-8 and np.allclose(x, img, atol=1e-5)
