"""adrt_amd.core.adrt_step / bdrt_step and their iterators (mirrors the
reference's tests/test_adrt_step.py, test_bdrt_step.py, test_adrt_iter.py and
test_bdrt_iter.py)."""

import itertools

import numpy as np
import pytest

import adrt_amd
from helpers import assert_bitwise, golden_small
from oracle import oracle as O

gpu = pytest.mark.gpu


@pytest.mark.parametrize("name", ["adrt_step", "bdrt_step"])
class TestStepCdefsValidation:
    def _fn(self, name):
        return getattr(adrt_amd._adrt_cdefs, name)

    def test_refuses_too_few_args(self, name):
        with pytest.raises(TypeError, match=rf"{name}.* expected 2 arguments, got 1"):
            self._fn(name)(np.zeros((4, 31, 16), dtype=np.float32))

    def test_refuses_too_many_args(self, name):
        with pytest.raises(TypeError, match=rf"{name}.* expected 2 arguments, got 3"):
            self._fn(name)(np.zeros((4, 31, 16)), 0, 0)

    def test_refuses_non_array(self, name):
        with pytest.raises(TypeError, match="must be a NumPy array or compatible subclass"):
            self._fn(name)(np.zeros((4, 31, 16)).tolist(), 0)

    def test_refuses_non_integer(self, name):
        with pytest.raises(TypeError, match="int"):
            self._fn(name)(np.zeros((4, 31, 16)), [])

    def test_refuses_out_of_range(self, name):
        arr = np.zeros((4, 31, 16), dtype=np.float32)
        with pytest.raises(ValueError, match=r"step \-1 is out of range.*adrt\.core\.num_iters"):
            self._fn(name)(arr, -1)
        with pytest.raises(ValueError, match=r"step 4 is out of range.*adrt\.core\.num_iters"):
            self._fn(name)(arr, 4)

    def test_refuses_int_overflow(self, name):
        with pytest.raises(OverflowError, match="Python int too large to convert to C int"):
            self._fn(name)(np.zeros((4, 31, 16)), 2**40)

    def test_public_refuses_float_step(self, name):
        with pytest.raises((TypeError, AttributeError)):
            getattr(adrt_amd.core, name)(np.zeros((4, 31, 16)), 0.0)

    def test_refuses_int32(self, name):
        with pytest.raises(TypeError, match="int32"):
            self._fn(name)(np.zeros((4, 31, 16), dtype=np.int32), 0)


def test_adrt_step_shape_message():
    with pytest.raises(ValueError, match="valid shape for ADRT, use adrt.core.adrt_init"):
        adrt_amd._adrt_cdefs.adrt_step(np.zeros((4, 30, 16)), 0)


def test_iter_validation_cpu():
    with pytest.raises(ValueError, match="must be square"):
        list(adrt_amd.core.adrt_iter(np.ones((16, 15), dtype=np.float32)))
    with pytest.raises(ValueError, match="between 2 and 3 dimensions, but had 4"):
        list(adrt_amd.core.adrt_iter(np.ones((2, 3, 16, 16), dtype=np.float32)))
    with pytest.raises(ValueError, match="must have a valid ADRT output shape"):
        list(adrt_amd.core.bdrt_iter(np.ones((4, 31, 15), dtype=np.float32)))
    with pytest.raises(ValueError, match="between 3 and 4 dimensions, but had 2"):
        list(adrt_amd.core.bdrt_iter(np.ones((31, 16), dtype=np.float32)))


@gpu
class TestStepCompute:
    @pytest.mark.parametrize("dt", ["float32", "float64"])
    @pytest.mark.parametrize("n", [2, 4, 8, 16])
    def test_golden_vectors(self, cuda, dt, n):
        g = golden_small()
        dat = g[f"{dt}_n{n}_dat"]
        for s in range(adrt_amd.core.num_iters(n)):
            assert_bitwise(adrt_amd.core.adrt_step(dat, s), g[f"{dt}_n{n}_adrt_step{s}"])
            assert_bitwise(adrt_amd.core.bdrt_step(dat, s), g[f"{dt}_n{n}_bdrt_step{s}"])

    @pytest.mark.parametrize("n", [64, 256])
    def test_parity_random(self, cuda, n):
        y = np.random.default_rng(n).standard_normal((2, 4, 2 * n - 1, n)).astype(np.float32)
        for s in range(adrt_amd.core.num_iters(n)):
            assert_bitwise(adrt_amd.core.adrt_step(y, s), O.adrt_step(y, s), f"adrt_step {s}")
            assert_bitwise(adrt_amd.core.bdrt_step(y, s), O.bdrt_step(y, s), f"bdrt_step {s}")

    @pytest.mark.parametrize("size", [2, 4, 8])
    def test_materialize_transposes(self, cuda, size):
        n = size
        full = 4 * n * (2 * n - 1)
        basis = np.eye(full, dtype=np.float32).reshape((full, 4, 2 * n - 1, n))
        L = adrt_amd.core.num_iters(n)
        for k, j in zip(range(L), reversed(range(L))):
            b = adrt_amd.core.bdrt_step(basis, k).reshape(full, full)
            a = adrt_amd.core.adrt_step(basis, j).reshape(full, full)
            assert np.array_equal(a, b.T)

    @pytest.mark.parametrize("dtype", ["float32", "float64"])
    def test_adrt_iter_matches_adrt(self, cuda, dtype):
        x = np.arange(16 * 16).reshape(16, 16).astype(dtype)
        items = list(adrt_amd.core.adrt_iter(x))
        assert len(items) == adrt_amd.core.num_iters(16) + 1
        assert_bitwise(items[-1], adrt_amd.adrt(x))
        assert np.all(items[0] == adrt_amd.core.adrt_init(x))

    def test_adrt_iter_pairs_are_steps(self, cuda):
        x = np.random.default_rng(1).random((3, 16, 16))
        for i, (a, b) in enumerate(itertools.pairwise(adrt_amd.core.adrt_iter(x))):
            assert_bitwise(adrt_amd.core.adrt_step(a, step=i), b)

    @pytest.mark.parametrize("size", [1, 2, 4, 8, 16])
    def test_iter_lengths(self, cuda, size):
        x = np.ones((size, size), dtype=np.float32)
        assert len(list(adrt_amd.core.adrt_iter(x))) == adrt_amd.core.num_iters(size) + 1
        y = adrt_amd.adrt(x)
        assert len(list(adrt_amd.core.bdrt_iter(y))) == adrt_amd.core.num_iters(size)

    def test_bdrt_iter_matches_bdrt(self, cuda):
        y = np.arange(4 * 31 * 16).reshape(4, 31, 16).astype(np.float64)
        last = list(adrt_amd.core.bdrt_iter(y))[-1]
        assert np.allclose(last, adrt_amd.bdrt(y))

    def test_iter_copy_semantics(self, cuda):
        x = np.ones((16, 16), dtype=np.float32)
        assert all(a.flags.writeable and a.base is None for a in adrt_amd.core.adrt_iter(x))
        assert all(not a.flags.writeable for a in adrt_amd.core.adrt_iter(x, copy=False))
        y = adrt_amd.adrt(x)
        assert all(a.flags.writeable and a.base is None for a in adrt_amd.core.bdrt_iter(y))
        assert all(not a.flags.writeable for a in adrt_amd.core.bdrt_iter(y, copy=False))

    def test_iter_refuses_int32(self, cuda):
        with pytest.raises(TypeError, match="int32"):
            list(adrt_amd.core.adrt_iter(np.ones((16, 16), dtype=np.int32)))


@pytest.mark.gpu
class TestStageChain:
    """adrt_hip_stage_chain: forward stages [first, last) in one call, equal
    to that many adrt_step calls (and to the oracle's steps)."""

    @pytest.mark.parametrize("n", [1, 2, 8, 64, 256])
    @pytest.mark.parametrize("dt", ["float32", "float64"])
    def test_full_chain_matches_steps(self, cuda, n, dt):
        import torch
        from adrt_amd import _kernels
        from oracle import oracle as O

        a = np.random.default_rng(n).standard_normal((2, 4, 2 * n - 1, n)).astype(dt)
        L = O.num_iters(n)
        want = a
        for k in range(L):
            want = O.adrt_step(want, k)
        t = torch.from_numpy(a).to(cuda)
        if L == 0:
            return
        got = _kernels.stage_chain(t)
        assert got.cpu().numpy().tobytes() == np.ascontiguousarray(want).tobytes()

    def test_partial_ranges(self, cuda):
        import torch
        from adrt_amd import _kernels

        a = torch.from_numpy(np.random.default_rng(1).standard_normal((3, 4, 127, 64))).to(cuda)
        for first, last in [(0, 1), (2, 3), (1, 5), (3, 6)]:
            want = a
            for k in range(first, last):
                want = _kernels.adrt_step(want, k)
            assert torch.equal(_kernels.stage_chain(a, first, last), want)

    def test_errors(self, cuda):
        import torch
        from adrt_amd import _lib

        lib = _lib.load()
        a = torch.zeros((1, 4, 15, 8), device=cuda)
        out = torch.empty_like(a)
        ws = torch.empty(lib.adrt_hip_stage_chain_workspace(_lib.F32, 1, 8), dtype=torch.uint8,
                         device=cuda)
        call = lambda f, l, w=ws: lib.adrt_hip_stage_chain(  # noqa: E731
            a.data_ptr(), out.data_ptr(), w.data_ptr(), w.numel(), _lib.F32, 1, 8, f, l,
            torch.cuda.current_stream().cuda_stream)
        assert call(0, 4) == _lib.ERR_STEP and call(2, 2) == _lib.ERR_STEP
        assert call(-1, 2) == _lib.ERR_STEP
        small = torch.empty(16, dtype=torch.uint8, device=cuda)
        assert call(0, 3, small) == _lib.ERR_WORKSPACE
        assert call(1, 2, small) == _lib.OK  # one step needs no workspace
