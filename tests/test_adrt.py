"""adrt_amd.adrt and adrt_amd._adrt_cdefs.adrt (mirrors the reference's
tests/test_adrt.py).  Validation runs on CPU; compute cases are GPU tests
checked bit for bit against the oracle and the brute force."""

import numpy as np
import pytest

import adrt_amd
from helpers import assert_bitwise, brute_force_adrt, canonical_sha256, golden_hashes, golden_small, make_unaligned
from oracle import oracle as O

gpu = pytest.mark.gpu


class TestAdrtCdefsValidation:
    def test_refuses_int32(self):
        with pytest.raises(TypeError, match="int32"):
            adrt_amd._adrt_cdefs.adrt(np.zeros((16, 16), dtype=np.int32))

    def test_refuses_non_square(self):
        with pytest.raises(ValueError, match="must be square"):
            adrt_amd._adrt_cdefs.adrt(np.zeros((16, 32), dtype=np.float32))

    def test_refuses_four_dim(self):
        with pytest.raises(ValueError, match="between 2 and 3 dimensions, but had 4"):
            adrt_amd._adrt_cdefs.adrt(np.zeros((5, 3, 16, 16), dtype=np.float32))

    def test_refuses_non_power_of_two(self):
        with pytest.raises(ValueError, match="power of two shape"):
            adrt_amd._adrt_cdefs.adrt(np.zeros((31, 31), dtype=np.float32))

    def test_refuses_non_array(self):
        with pytest.raises(TypeError, match="must be a NumPy array or compatible subclass"):
            adrt_amd._adrt_cdefs.adrt(None)
        with pytest.raises(TypeError, match="must be a NumPy array or compatible subclass"):
            adrt_amd._adrt_cdefs.adrt([[1.0, 2.0], [3.0, 4.0]])

    def test_refuses_fortran_order(self):
        with pytest.raises(ValueError, match=r"must be .*C-order"):
            adrt_amd._adrt_cdefs.adrt(np.zeros((32, 32), dtype=np.float32, order="F"))

    def test_refuses_c_non_contiguous(self):
        a = np.zeros((64, 32), dtype=np.float32, order="F")[::2]
        assert a.shape == (32, 32) and not a.flags["C_CONTIGUOUS"]
        with pytest.raises(ValueError, match=r"must be .*contiguous"):
            adrt_amd._adrt_cdefs.adrt(a)

    def test_refuses_byteswapped(self):
        a = np.ones((16, 16), dtype=np.dtype(np.float32).newbyteorder())
        with pytest.raises(ValueError, match=r"must be .*native byte order"):
            adrt_amd._adrt_cdefs.adrt(a)

    def test_refuses_zero_axis_array(self):
        with pytest.raises(ValueError, match="found zero in dimension 0"):
            adrt_amd._adrt_cdefs.adrt(np.zeros((0, 32, 32), dtype=np.float32))

    def test_refuses_zero_size_planes(self):
        with pytest.raises(ValueError, match=r"found zero in dimension (0|1)"):
            adrt_amd._adrt_cdefs.adrt(np.zeros((0, 0), dtype=np.float32))

    @pytest.mark.parametrize("dtype", ["float32", "float64"])
    def test_refuses_unaligned(self, dtype):
        with pytest.raises(ValueError, match=r"must be .*aligned"):
            adrt_amd._adrt_cdefs.adrt(make_unaligned(np.zeros((16, 16), dtype=dtype)))

    def test_dtype_checked_after_shape(self):
        # adrt_cdefs_py.cpp:282-298: shape errors win over dtype errors.
        with pytest.raises(ValueError, match="power of two"):
            adrt_amd._adrt_cdefs.adrt(np.zeros((3, 3), dtype=np.int32))

    def test_public_refuses_non_array(self):
        with pytest.raises(TypeError, match=r"array must be numpy\.ndarray, but got list"):
            adrt_amd.adrt([[1.0]])

    def test_public_refuses_int32(self):
        with pytest.raises(TypeError, match="int32"):
            adrt_amd.adrt(np.zeros((16, 16), dtype=np.int32))


@gpu
class TestAdrtCompute:
    def test_accepts_float32_three_dim(self, cuda):
        out = adrt_amd._adrt_cdefs.adrt(np.zeros((5, 16, 16), dtype=np.float32))
        assert out.shape == (5, 4, 31, 16) and out.dtype == np.float32

    @pytest.mark.parametrize("dtype", ["float32", "float64"])
    def test_returned_dtype(self, cuda, dtype):
        assert adrt_amd.adrt(np.zeros((16, 16), dtype=dtype)).dtype == np.dtype(dtype)

    def test_accepts_fortran_order(self, cuda):
        a = np.asfortranarray(np.random.default_rng(0).random((16, 16)).astype(np.float32))
        assert_bitwise(adrt_amd.adrt(a), O.adrt(np.ascontiguousarray(a)))

    def test_accepts_c_non_contiguous(self, cuda):
        a = np.random.default_rng(1).random((64, 32)).astype(np.float32)[::2]
        assert_bitwise(adrt_amd.adrt(a), O.adrt(np.ascontiguousarray(a)))

    @pytest.mark.parametrize("dtype", ["float32", "float64"])
    def test_accepts_unaligned(self, cuda, dtype):
        base = np.arange(16 * 16, dtype=dtype).reshape((16, 16))
        assert np.all(adrt_amd.adrt(base) == adrt_amd.adrt(make_unaligned(base)))

    def test_accepts_byteswapped(self, cuda):
        a = np.random.default_rng(2).random((16, 16)).astype(np.float32)
        sw = a.astype(a.dtype.newbyteorder("S"))
        assert np.all(adrt_amd.adrt(a) == adrt_amd.adrt(sw))

    @pytest.mark.parametrize("n", [1, 2, 4, 8, 16])
    def test_brute_force_integer(self, cuda, n):
        img = np.random.default_rng(n).integers(-20, 20, size=(n, n)).astype(np.float64)
        out = adrt_amd.adrt(img)
        assert np.array_equal(out, brute_force_adrt(img))
        m = n
        for q in range(4):
            assert np.all(np.tril(out[q, -m:], k=-1) == 0)

    def test_all_zeros(self, cuda):
        out = adrt_amd.adrt(np.zeros((32, 32)))
        assert out.shape == (4, 63, 32) and np.all(out == 0)

    def test_small_1x1_batch(self, cuda):
        a = np.arange(5, dtype="float64").reshape((5, 1, 1)) + 1
        assert np.array_equal(adrt_amd.adrt(a), np.stack([a] * 4, axis=1))

    @pytest.mark.parametrize("dt", ["float32", "float64"])
    @pytest.mark.parametrize("n", [1, 2, 4, 8, 16])
    def test_golden_vectors(self, cuda, dt, n):
        g = golden_small()
        assert_bitwise(adrt_amd.adrt(g[f"{dt}_n{n}_img"]), g[f"{dt}_n{n}_adrt"])

    def test_parity_config_8x128_seed0(self, cuda):
        """Forward parity case: 8 x 128^2 float32, U[0,1) from seed 0."""
        x = np.random.default_rng(0).random((8, 128, 128), dtype=np.float32)
        out = adrt_amd.adrt(x)
        assert out.shape == (8, 4, 255, 128)
        assert canonical_sha256(out) == golden_hashes()["adrt_f32_8x128_seed0"]
        assert_bitwise(out, O.adrt(x))

    @pytest.mark.parametrize("dt", ["float32", "float64"])
    @pytest.mark.parametrize("n", [32, 64, 128, 256, 512, 1024])
    def test_parity_random(self, cuda, dt, n):
        rng = np.random.default_rng(n)
        b = max(1, 2048 // n)
        x = (rng.standard_normal((b, n, n)) * 100).astype(dt)
        x[0, 0, :3] = [-0.0, np.inf, -np.inf]
        x[-1, -1, -1] = np.nan
        assert_bitwise(adrt_amd.adrt(x), O.adrt(x), f"adrt {dt} n={n}")

    @pytest.mark.parametrize("dt", ["float32", "float64"])
    def test_parity_2048(self, cuda, dt):
        x = np.random.default_rng(2048).standard_normal((2, 2048, 2048)).astype(dt)
        assert_bitwise(adrt_amd.adrt(x), O.adrt(x), f"adrt {dt} n=2048")

    def test_parity_4096(self, cuda):
        x = np.random.default_rng(4096).standard_normal((1, 4096, 4096)).astype(np.float32)
        assert_bitwise(adrt_amd.adrt(x), O.adrt(x), "adrt n=4096")

    def test_parity_1024_hash(self, cuda):
        big = np.random.default_rng(1).random((1, 1024, 1024), dtype=np.float32)
        assert canonical_sha256(adrt_amd.adrt(big)) == golden_hashes()["adrt_f32_1x1024_seed1"]

    def test_torch_device_tensor(self, cuda):
        import torch

        x = torch.rand(3, 64, 64, device=cuda, dtype=torch.float32)
        y = adrt_amd.adrt(x)
        assert y.is_cuda and y.shape == (3, 4, 127, 64)
        assert_bitwise(y.cpu().numpy(), O.adrt(x.cpu().numpy()))
