"""adrt_amd.iadrt (mirrors the reference's tests/test_iadrt.py)."""

import numpy as np
import pytest

import adrt_amd
from helpers import assert_bitwise, canonical_sha256, golden_hashes, golden_small
from oracle import oracle as O

gpu = pytest.mark.gpu


class TestIAdrtCdefsValidation:
    def test_refuses_int32(self):
        with pytest.raises(TypeError, match="int32"):
            adrt_amd._adrt_cdefs.iadrt(np.zeros((4, 31, 16), dtype=np.int32))

    def test_refuses_mismatched_shape(self):
        for shape in [(4, 31, 15), (4, 30, 16)]:
            with pytest.raises(ValueError, match="must have a valid ADRT output shape"):
                adrt_amd._adrt_cdefs.iadrt(np.zeros(shape, dtype=np.float32))

    def test_refuses_five_dim(self):
        with pytest.raises(ValueError, match="between 3 and 4 dimensions, but had 5"):
            adrt_amd._adrt_cdefs.iadrt(np.zeros((6, 5, 4, 31, 16), dtype=np.float32))

    def test_refuses_non_array(self):
        with pytest.raises(TypeError, match="must be a NumPy array or compatible subclass"):
            adrt_amd._adrt_cdefs.iadrt([[[1.0]]])

    def test_refuses_c_non_contiguous(self):
        a = np.zeros((4, 31, 32), dtype=np.float32, order="F")[:, :, ::2]
        with pytest.raises(ValueError, match=r"must be .*contiguous"):
            adrt_amd._adrt_cdefs.iadrt(a)

    def test_refuses_zero_size_planes(self):
        with pytest.raises(ValueError, match=r"found zero in dimension (1|2)"):
            adrt_amd._adrt_cdefs.iadrt(np.zeros((4, 0, 0), dtype=np.float32))


@gpu
class TestIAdrtCompute:
    @pytest.mark.parametrize("dtype", ["float32", "float64"])
    def test_returned_dtype(self, cuda, dtype):
        assert adrt_amd.iadrt(np.zeros((4, 31, 16), dtype=dtype)).dtype == np.dtype(dtype)

    @pytest.mark.parametrize("size", [16, 32])
    def test_unique_values_round_trip(self, cuda, size):
        a = np.arange(size**2).reshape((size, size)).astype("float32")
        inv = np.mean(adrt_amd.utils.truncate(adrt_amd.iadrt(adrt_amd.adrt(a))), axis=0)
        assert inv.dtype == a.dtype and np.allclose(inv, a)

    def test_batch_round_trip(self, cuda):
        a = np.arange(4 * 32 * 32).reshape((4, 32, 32)).astype("float32")
        inv = np.mean(adrt_amd.utils.truncate(adrt_amd.iadrt(adrt_amd.adrt(a))), axis=1)
        assert np.allclose(inv, a)

    def test_small_1x1_batch(self, cuda):
        e = np.arange(5, dtype="float64").reshape((5, 1, 1)) + 1
        out = np.mean(adrt_amd.utils.truncate(adrt_amd.iadrt(np.stack([e] * 4, axis=1))), axis=1)
        assert np.allclose(out, e)

    @pytest.mark.parametrize("dt", ["float32", "float64"])
    @pytest.mark.parametrize("n", [1, 2, 4, 8, 16])
    def test_golden_vectors(self, cuda, dt, n):
        g = golden_small()
        assert_bitwise(adrt_amd.iadrt(g[f"{dt}_n{n}_dat"]), g[f"{dt}_n{n}_iadrt"])

    @pytest.mark.parametrize("dt", ["float32", "float64"])
    @pytest.mark.parametrize("n", [32, 128, 512])
    def test_parity_random(self, cuda, dt, n):
        rng = np.random.default_rng(n + 2)
        b = max(1, 512 // n)
        y = rng.standard_normal((b, 4, 2 * n - 1, n)).astype(dt)
        assert_bitwise(adrt_amd.iadrt(y), O.iadrt(y), f"iadrt {dt} n={n}")

    @pytest.mark.parametrize("dt", ["float32", "float64"])
    @pytest.mark.parametrize("n", [2, 8, 64, 256, 2048])
    def test_every_variant_bitwise(self, cuda, dt, n):
        """The fused tail (default and half length) and the one-kernel-per-
        stage path give the oracle's bytes, on device tensors."""
        import torch
        from adrt_amd import _kernels, _lib, batched

        rng = np.random.default_rng(n + 11)
        b = 1 if n == 2048 else 3
        y = rng.standard_normal((b, 4, 2 * n - 1, n)).astype(dt)
        y.flat[:2] = [-0.0, 0.0]
        want = O.iadrt(y)
        yd = torch.from_numpy(y).to(cuda)
        outs = []
        try:
            for v in (_lib.VARIANT_DEFAULT, _lib.VARIANT_ALT_SPLIT, _lib.VARIANT_STREAMING):
                batched.set_config(_lib.OP_IADRT, n, v)
                outs.append(_kernels.iadrt(yd).cpu().numpy())
        finally:
            batched.set_config(_lib.OP_IADRT, n, _lib.VARIANT_DEFAULT)
        for o in outs[1:]:
            assert o.tobytes() == outs[0].tobytes()
        assert_bitwise(outs[0], want, f"iadrt {dt} n={n}")

    def test_integer_round_trip_exact(self, cuda):
        x = np.random.default_rng(5).integers(-9, 9, size=(3, 64, 64)).astype(np.float64)
        inv = adrt_amd.utils.truncate(adrt_amd.iadrt(adrt_amd.adrt(x))).mean(axis=-3)
        assert np.array_equal(inv, x)

    def test_parity_hash(self, cuda):
        x = np.random.default_rng(0).random((8, 128, 128), dtype=np.float32)
        out = adrt_amd.iadrt(O.adrt(x))
        assert canonical_sha256(out) == golden_hashes()["iadrt_f32_of_adrt_8x128_seed0"]
