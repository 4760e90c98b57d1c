"""adrt_init, truncate, num_iters, threading_enabled (mirrors the reference's
tests/test_adrt_init.py, test_truncate.py, test_num_iters.py and
test_threading_enabled.py)."""

import math
import re

import numpy as np
import pytest

import adrt_amd
from oracle import oracle as O

gpu = pytest.mark.gpu
INT_FLOAT = ["float32", "float64", "int32", "int64"]


# ---- num_iters (pure host) ----------------------------------------------------


@pytest.mark.parametrize("v", [-1, -2])
def test_num_iters_rejects_negative(v):
    with pytest.raises(ValueError, match=rf"non-negative value required.*got {re.escape(str(v))}"):
        adrt_amd.core.num_iters(v)


@pytest.mark.parametrize("v", [4.0, None, (), []])
def test_num_iters_rejects_non_integer(v):
    with pytest.raises(TypeError):
        adrt_amd.core.num_iters(v)


@pytest.mark.parametrize("v", range(0, 17))
def test_num_iters_values(v):
    expect = 0 if v <= 1 else math.ceil(math.log2(v))
    assert adrt_amd.core.num_iters(v) == expect
    assert adrt_amd.core.num_iters(v) == O.num_iters(v)


def test_num_iters_numpy_and_bool():
    for cls in (np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64):
        assert adrt_amd.core.num_iters(cls(20)) == 5
    assert adrt_amd.core.num_iters(True) == 0
    with pytest.raises(TypeError):
        adrt_amd.core.num_iters(np.float32(4))


@pytest.mark.parametrize("exp", [32, 64, 128])
@pytest.mark.parametrize("off", [-1, 0, 1])
def test_num_iters_large(exp, off):
    assert adrt_amd.core.num_iters(2**exp + off) == exp + (1 if off > 0 else 0)


def test_threading_enabled():
    assert isinstance(adrt_amd.core.threading_enabled(), bool)
    assert adrt_amd.core.threading_enabled() == adrt_amd._adrt_cdefs.OPENMP_ENABLED


# ---- adrt_init / truncate validation (host) ----------------------------------


def test_init_refuses_non_array():
    with pytest.raises(TypeError, match=r"must be numpy\.ndarray"):
        adrt_amd.core.adrt_init(None)
    with pytest.raises(TypeError, match=r"must be numpy\.ndarray"):
        adrt_amd.core.adrt_init([[1.0, 2.0], [3.0, 4.0]])


def test_init_refuses_bad_shapes():
    with pytest.raises(ValueError, match="between 2 and 3 dimensions, but had 4"):
        adrt_amd.core.adrt_init(np.ones((2, 3, 16, 16), dtype=np.float32))
    with pytest.raises(ValueError, match="between 2 and 3 dimensions, but had 1"):
        adrt_amd.core.adrt_init(np.ones(16, dtype=np.float32))
    with pytest.raises(ValueError, match="must be square"):
        adrt_amd.core.adrt_init(np.ones((3, 16, 15), dtype=np.float32))
    with pytest.raises(ValueError, match="power of two shape"):
        adrt_amd.core.adrt_init(np.ones((2, 7, 7), dtype=np.float32))
    with pytest.raises(ValueError, match="found zero in dimension 0"):
        adrt_amd.core.adrt_init(np.zeros((0, 32, 32), dtype=np.float32))


@pytest.mark.parametrize("shape", [(4, 0, 0), (4, 4, 2), (7, 4)])
def test_truncate_rejects_invalid(shape):
    with pytest.raises(ValueError, match=r"unsuitable shape .*ADRT output"):
        adrt_amd.utils.truncate(np.ones(shape, dtype=np.float32))


# ---- device ---------------------------------------------------------------------


@gpu
@pytest.mark.parametrize("dtype", INT_FLOAT + ["int8", "uint16", "bool", "complex64"])
def test_init_matches_oracle(cuda, dtype):
    x = (np.arange(3 * 16 * 16) % 7).reshape(3, 16, 16).astype(dtype)
    out = adrt_amd.core.adrt_init(x)
    assert out.dtype == x.dtype and out.shape == (3, 4, 31, 16)
    assert np.array_equal(out, O.adrt_init(x))
    assert np.array_equal(adrt_amd.core.adrt_init(x[0]), O.adrt_init(x[0]))


@gpu
@pytest.mark.parametrize("dtype", INT_FLOAT)
def test_truncate_inverts_init(cuda, dtype):
    x = np.arange(5 * 16 * 16).reshape(5, 16, 16).astype(dtype)
    out = adrt_amd.utils.truncate(adrt_amd.core.adrt_init(x))
    assert out.shape == (5, 4, 16, 16) and out.dtype == x.dtype
    assert np.array_equal(np.expand_dims(x, 1).repeat(4, axis=1), out)


@gpu
def test_truncate_matches_oracle_any_leading_dims(cuda):
    y = np.random.default_rng(0).standard_normal((2, 3, 4, 11, 6)).astype(np.float32)
    assert np.array_equal(adrt_amd.utils.truncate(y), O.truncate(y))
    z = np.random.default_rng(1).standard_normal((4, 9, 5))  # non power of two N
    assert np.array_equal(adrt_amd.utils.truncate(z), O.truncate(z))


@gpu
def test_truncate_accepts_bdrt_and_iadrt(cuda):
    y = np.ones((4, 31, 16))
    assert adrt_amd.utils.truncate(adrt_amd.bdrt(y)).shape == (4, 16, 16)
    assert adrt_amd.utils.truncate(adrt_amd.iadrt(y)).shape == (4, 16, 16)


@gpu
def test_init_torch_tensor(cuda):
    import torch

    x = torch.arange(2 * 8 * 8, device=cuda, dtype=torch.int64).reshape(2, 8, 8)
    out = adrt_amd.core.adrt_init(x)
    assert out.is_cuda and np.array_equal(out.cpu().numpy(), O.adrt_init(x.cpu().numpy()))
