"""Multigrid operators and the FMG inverse (mirrors the reference's
tests/test_press_fmg_*.py, test_iadrt_fmg.py, test_iadrt_fmg_step.py and
test_iadrt_fmg_iter.py)."""

import contextlib
import functools
import itertools

import numpy as np
import pytest

import adrt_amd
from adrt_amd import _wrappers as W
from helpers import assert_bitwise, golden_small
from oracle import oracle as O

gpu = pytest.mark.gpu
DT = ["float32", "float64"]


# ---- validation (host) ----------------------------------------------------------


def test_restriction_rejects():
    for shape in [(4, 5, 3), (4, 10, 6), (2, 11, 6)]:
        with pytest.raises(ValueError, match="must have a valid ADRT output shape"):
            W._press_fmg_restriction(np.zeros(shape, dtype=np.float32))
    with pytest.raises(TypeError, match="int32"):
        W._press_fmg_restriction(np.zeros((4, 11, 6), dtype=np.int32))


def test_prolongation_rejects():
    with pytest.raises(ValueError, match="between 2 and 3 dimensions, but had 1"):
        W._press_fmg_prolongation(np.arange(7).astype("float32"))
    with pytest.raises(ValueError, match="between 2 and 3 dimensions, but had 4"):
        W._press_fmg_prolongation(np.zeros((2, 3, 5, 6), dtype="float32"))
    with pytest.raises(TypeError, match="int32"):
        W._press_fmg_prolongation(np.zeros((5, 6), dtype="int32"))


def test_highpass_rejects():
    a = np.ones((1, 2), dtype=np.float32)
    with pytest.raises(ValueError, match="array is too small"):
        W._press_fmg_highpass(a)
    with pytest.raises(ValueError, match="array is too small"):
        W._press_fmg_highpass(np.ascontiguousarray(a.T))
    with pytest.raises(TypeError, match="int32"):
        W._press_fmg_highpass(np.ones((16, 16), dtype=np.int32))


def test_iadrt_fmg_rejects():
    z = np.zeros((4, 15, 8), dtype="float32")
    with pytest.raises(ValueError, match=r"must be.*int"):
        adrt_amd.iadrt_fmg(z, max_iters=2.0)
    with pytest.raises(ValueError, match=r"must allow at least one iteration.*specified 0"):
        adrt_amd.iadrt_fmg(z, max_iters=0)
    with pytest.raises(ValueError, match=r"batch dimension not supported.*got 4 dimensions"):
        adrt_amd.iadrt_fmg(np.zeros((3, 4, 15, 8), dtype="float32"), max_iters=50)


def test_fmg_iter_rejects_dims():
    with pytest.raises(ValueError, match="between 3 and 4 dimensions, but had 5"):
        next(adrt_amd.core.iadrt_fmg_iter(np.ones((2, 3, 4, 31, 16), dtype="float32")))
    with pytest.raises(ValueError, match="between 3 and 4 dimensions, but had 2"):
        next(adrt_amd.core.iadrt_fmg_iter(np.ones((31, 16), dtype="float32")))


# ---- device ---------------------------------------------------------------------


@gpu
@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_operators_golden(cuda, dt, n):
    g = golden_small()
    k = f"{dt}_n{n}"
    assert_bitwise(W._press_fmg_restriction(g[f"{k}_dat"]), g[f"{k}_restriction"])
    assert_bitwise(W._press_fmg_highpass(g[f"{k}_hp_in"]), g[f"{k}_highpass"])
    assert_bitwise(W._press_fmg_prolongation(g[f"{k}_hp_in"]), g[f"{k}_prolongation"])


@gpu
@pytest.mark.parametrize("dt", DT)
def test_operators_parity_random(cuda, dt):
    rng = np.random.default_rng(3)
    a = rng.standard_normal((3, 4, 255, 128)).astype(dt)
    assert_bitwise(W._press_fmg_restriction(a), O.fmg_restriction(a))
    b = rng.standard_normal((3, 37, 53)).astype(dt)
    assert_bitwise(W._press_fmg_highpass(b), O.fmg_highpass(b))
    assert_bitwise(W._press_fmg_prolongation(b), O.fmg_prolongation(b))
    assert_bitwise(W._press_fmg_highpass(b[0]), O.fmg_highpass(b[0]))


@gpu
@pytest.mark.parametrize("dt", DT)
def test_highpass_fixed_points(cuda, dt):
    x, y = np.meshgrid(np.arange(7), np.arange(10))
    cb = ((x % 2) ^ (y % 2)).astype(dt) * 2 - 1
    assert np.allclose(W._press_fmg_highpass(cb), cb)
    assert np.allclose(W._press_fmg_highpass(np.full((14, 5), 5, dtype=dt)), 0)


@gpu
def test_prolongation_tiny(cuda):
    assert np.array_equal(W._press_fmg_prolongation(np.array([[3.0]])), np.full((2, 2), 3.0))


@gpu
@pytest.mark.parametrize("dt", DT)
def test_fmg_step_bitwise_vs_oracle(cuda, dt):
    arr = np.arange(4 * 31 * 16).reshape((4, 31, 16)).astype(dt)
    assert_bitwise(adrt_amd.core.iadrt_fmg_step(arr), O.iadrt_fmg_step(arr))
    rng = np.random.default_rng(8)
    y = O.adrt(rng.random((3, 64, 64)).astype(dt))
    assert_bitwise(adrt_amd.core.iadrt_fmg_step(y), O.iadrt_fmg_step(y))


@gpu
@pytest.mark.parametrize("dt", DT)
def test_fmg_step_batch_shape(cuda, dt):
    arr = np.arange(3 * 4 * 31 * 16).reshape((3, 4, 31, 16)).astype(dt)
    out = adrt_amd.core.iadrt_fmg_step(arr)
    assert out.dtype == arr.dtype and out.shape == (3, 16, 16)
    ref = np.stack([O.iadrt_fmg_step(arr[i]) for i in range(3)])
    assert np.allclose(out, ref, atol=1e-3)


@gpu
@pytest.mark.parametrize("dt", DT)
def test_iadrt_fmg_unique_values(cuda, dt):
    orig = np.arange(64).reshape((8, 8)).astype(dt)
    inv = adrt_amd.iadrt_fmg(adrt_amd.adrt(orig), max_iters=50)
    assert inv.dtype == orig.dtype and inv.shape == orig.shape and inv.flags.writeable
    assert np.allclose(inv, orig, atol=1e-3)


@pytest.fixture
def counting_fmg_step(monkeypatch):
    count = 0
    orig = adrt_amd.core.iadrt_fmg_step

    @functools.wraps(orig)
    def counting(*args, **kwargs):
        nonlocal count
        count += 1
        return orig(*args, **kwargs)

    monkeypatch.setattr(adrt_amd.core, "iadrt_fmg_step", counting)
    return lambda: count


@gpu
@pytest.mark.parametrize("val", [0, np.nan, np.inf, -np.inf])
def test_stops_quickly_on_edge_case(cuda, counting_fmg_step, val):
    a = np.full((4, 15, 8), val, dtype=np.float32)
    with np.errstate(invalid="ignore") if np.isinf(val) else contextlib.nullcontext():
        inv = adrt_amd.iadrt_fmg(a, max_iters=50)
    assert counting_fmg_step() <= 2
    assert inv.shape == (8, 8) and inv.dtype == a.dtype
    assert np.isinf(val) or np.allclose(inv, val, equal_nan=True)


@gpu
@pytest.mark.parametrize("max_iters", [1, 2, 3, 4])
def test_max_iters_limits(cuda, counting_fmg_step, max_iters):
    orig = np.arange(256).reshape((16, 16)).astype("float64")
    adrt_amd.iadrt_fmg(adrt_amd.adrt(orig), max_iters=max_iters)
    assert counting_fmg_step() == max_iters


@gpu
def test_iadrt_fmg_matches_oracle_iterate(cuda):
    orig = np.random.default_rng(4).random((32, 32))
    a = O.adrt(orig)
    assert_bitwise(adrt_amd.iadrt_fmg(a, max_iters=3), O.iadrt_fmg(a, max_iters=3))


@gpu
def test_fmg_iter_copy_semantics_and_monotone(cuda):
    z = np.zeros((4, 31, 16), dtype="float32")
    assert all(a.flags.writeable and a.base is None
               for a in itertools.islice(adrt_amd.core.iadrt_fmg_iter(z), 5))
    assert all(not a.flags.writeable
               for a in itertools.islice(adrt_amd.core.iadrt_fmg_iter(z, copy=False), 5))
    steps = np.linspace(-10, 10, 32)
    x, y = np.meshgrid(steps, steps)
    img = np.exp(-np.hypot(x, y))
    res = np.linalg.norm(
        img[None] - np.array(list(itertools.islice(adrt_amd.core.iadrt_fmg_iter(adrt_amd.adrt(img)), 10))),
        axis=(-1, -2))
    assert np.all(np.diff(res) <= 0)


@gpu
@pytest.mark.parametrize("dt", DT)
def test_fmg_iter_batch(cuda, dt):
    a = np.arange(3 * 4 * 31 * 16).reshape((3, 4, 31, 16)).astype(dt)
    its = list(itertools.islice(adrt_amd.core.iadrt_fmg_iter(a), 3))
    ref = list(itertools.islice(O.iadrt_fmg_iter(a), 3))
    for x, r in zip(its, ref):
        assert_bitwise(x, r)


def test_iadrt_fmg_batched_rejects():
    with pytest.raises(ValueError, match="4 dimensions"):
        adrt_amd.core.iadrt_fmg_batched(np.zeros((4, 15, 8), dtype="float32"))
    with pytest.raises(ValueError, match="at least one iteration"):
        adrt_amd.core.iadrt_fmg_batched(np.zeros((2, 4, 15, 8), dtype="float32"), max_iters=0)
    with pytest.raises(ValueError, match="valid ADRT output shape"):
        adrt_amd.core.iadrt_fmg_batched(np.zeros((2, 4, 14, 8), dtype="float32"))


@gpu
@pytest.mark.parametrize("dt", ["float32", "float64"])
@pytest.mark.parametrize("n,max_iters", [(16, None), (64, None), (64, 2), (8, 1)])
def test_iadrt_fmg_batched_equals_per_image(cuda, dt, n, max_iters):
    """Each image of the batched FMG inverse is the unbatched iadrt_fmg of
    that image, bit for bit (its own stopping point included); images that
    stop at different iterations are mixed in one batch."""
    rng = np.random.default_rng(n)
    imgs = rng.random((4, n, n)).astype(dt)
    imgs[1] = 0.5                              # a flat image stops at once
    imgs[2] = np.round(imgs[2] * 8)            # an integer image
    a = adrt_amd.adrt(imgs)
    a[3] += rng.standard_normal(a[3].shape).astype(dt) * 0.01  # inconsistent data
    got = adrt_amd.core.iadrt_fmg_batched(a, max_iters=max_iters)
    assert got.shape == (4, n, n) and got.dtype == np.dtype(dt)
    for b in range(4):
        assert_bitwise(got[b], adrt_amd.iadrt_fmg(a[b], max_iters=max_iters), f"image {b}")


@gpu
def test_iadrt_fmg_batched_device_tensors(cuda):
    """Device tensors in, device tensor out, same values as NumPy in."""
    import torch

    imgs = np.random.default_rng(3).random((3, 32, 32)).astype("float32")
    a = adrt_amd.adrt(imgs)
    got = adrt_amd.core.iadrt_fmg_batched(torch.from_numpy(a).to(cuda))
    assert got.is_cuda and tuple(got.shape) == (3, 32, 32)
    assert_bitwise(got.cpu().numpy(), adrt_amd.core.iadrt_fmg_batched(a))
