"""adrt_amd.bdrt (mirrors the reference's tests/test_bdrt.py)."""

import numpy as np
import pytest

import adrt_amd
from helpers import assert_bitwise, canonical_sha256, digital_line, golden_hashes, golden_small
from oracle import oracle as O

gpu = pytest.mark.gpu


class TestBdrtCdefsValidation:
    def test_refuses_int32(self):
        with pytest.raises(TypeError, match="int32"):
            adrt_amd._adrt_cdefs.bdrt(np.zeros((4, 31, 16), dtype=np.int32))

    def test_refuses_mismatched_shape(self):
        for shape in [(4, 31, 15), (4, 30, 16)]:
            with pytest.raises(ValueError, match="must have a valid ADRT output shape"):
                adrt_amd._adrt_cdefs.bdrt(np.zeros(shape, dtype=np.float32))

    def test_refuses_five_dim(self):
        with pytest.raises(ValueError, match="between 3 and 4 dimensions, but had 5"):
            adrt_amd._adrt_cdefs.bdrt(np.zeros((6, 5, 4, 31, 16), dtype=np.float32))

    def test_refuses_non_power_of_two(self):
        with pytest.raises(ValueError, match="must have a valid ADRT output shape"):
            adrt_amd._adrt_cdefs.bdrt(np.zeros((4, 33, 17), dtype=np.float32))

    def test_refuses_non_array(self):
        with pytest.raises(TypeError, match="must be a NumPy array or compatible subclass"):
            adrt_amd._adrt_cdefs.bdrt(None)

    def test_refuses_fortran_order(self):
        with pytest.raises(ValueError, match=r"must be .*C-order"):
            adrt_amd._adrt_cdefs.bdrt(np.zeros((4, 31, 16), dtype=np.float32, order="F"))

    def test_refuses_byteswapped(self):
        a = np.ones((4, 31, 16), dtype=np.dtype(np.float32).newbyteorder())
        with pytest.raises(ValueError, match=r"must be .*native byte order"):
            adrt_amd._adrt_cdefs.bdrt(a)

    def test_refuses_zero_axis_array(self):
        with pytest.raises(ValueError, match="found zero in dimension 0"):
            adrt_amd._adrt_cdefs.bdrt(np.zeros((0, 4, 31, 16), dtype=np.float32))

    def test_refuses_zero_size_planes(self):
        with pytest.raises(ValueError, match=r"found zero in dimension (1|2)"):
            adrt_amd._adrt_cdefs.bdrt(np.zeros((4, 0, 0), dtype=np.float32))

    def test_public_refuses_int32(self):
        with pytest.raises(TypeError, match="int32"):
            adrt_amd.bdrt(np.zeros((4, 31, 16), dtype=np.int32))


def _bdrt_sq(x):
    return np.mean(adrt_amd.utils.truncate(adrt_amd.bdrt(x)), axis=-3)


@gpu
class TestBdrtCompute:
    @pytest.mark.parametrize("dtype", ["float32", "float64"])
    def test_returned_dtype(self, cuda, dtype):
        assert adrt_amd.bdrt(np.zeros((4, 31, 16), dtype=dtype)).dtype == np.dtype(dtype)

    def test_accepts_four_dim(self, cuda):
        assert adrt_amd._adrt_cdefs.bdrt(np.zeros((5, 4, 31, 16), dtype=np.float32)).shape == (5, 4, 31, 16)

    def test_accepts_c_non_contiguous(self, cuda):
        a = np.random.default_rng(0).random((4, 62, 16)).astype(np.float32)[:, ::2]
        assert_bitwise(adrt_amd.bdrt(a), O.bdrt(np.ascontiguousarray(a)))

    def test_all_zeros(self, cuda):
        out = adrt_amd.bdrt(adrt_amd.adrt(np.zeros((16, 16), dtype=np.float32)))
        assert np.all(out == 0)

    def test_digital_line(self, cuda):
        size = 16
        sq = _bdrt_sq(_delta(size, 0, 4, 3))
        expect = np.zeros((size, size))
        for i, j in digital_line(size, size - 5, 3):
            expect[i, j] = 0.25
        assert np.allclose(sq, expect)

    def test_backprojected_delta_levels(self, cuda):
        size = 16
        a = np.zeros((size, size))
        a[size // 2, size // 2] = 1.0
        sq = _bdrt_sq(adrt_amd.adrt(a))
        assert np.allclose(1.0 * (sq == 16.0), a)
        assert np.sum(sq == 4) == 8 and np.sum(sq == 2) == 19

    def test_bdrt_zeros(self, cuda):
        size = 16
        out = adrt_amd.bdrt(adrt_amd.adrt(np.ones((size, size))))
        for i in range(4):
            assert np.linalg.norm(np.tril(out[i, size:, ::-1]), ord="fro") == 0.0

    def test_small_1x1_batch(self, cuda):
        e = np.arange(5, dtype="float64").reshape((5, 1, 1)) + 1
        assert np.allclose(_bdrt_sq(np.stack([e] * 4, axis=1)), e)

    @pytest.mark.parametrize("n", [1, 2, 4, 8])
    def test_materialize_adjoint(self, cuda, n):
        full, size = n * n, 4 * n * (2 * n - 1)
        a_mat = adrt_amd.adrt(np.eye(full, dtype=np.float32).reshape(full, n, n)).reshape(full, -1).T
        b_out = adrt_amd.bdrt(np.eye(size, dtype=np.float32).reshape(size, 4, 2 * n - 1, n))
        b_mat = np.sum(adrt_amd.utils.truncate(b_out), axis=1).reshape(size, full)
        assert np.array_equal(a_mat, b_mat)

    @pytest.mark.parametrize("dt", ["float32", "float64"])
    @pytest.mark.parametrize("n", [1, 2, 4, 8, 16])
    def test_golden_vectors(self, cuda, dt, n):
        g = golden_small()
        assert_bitwise(adrt_amd.bdrt(g[f"{dt}_n{n}_dat"]), g[f"{dt}_n{n}_bdrt"])

    @pytest.mark.parametrize("dt", ["float32", "float64"])
    @pytest.mark.parametrize("n", [32, 64, 128, 256, 512])
    def test_parity_random(self, cuda, dt, n):
        rng = np.random.default_rng(n + 1)
        b = max(1, 1024 // n)
        y = (rng.standard_normal((b, 4, 2 * n - 1, n)) * 10).astype(dt)
        y[0, 0, 0, :2] = [-0.0, np.inf]
        assert_bitwise(adrt_amd.bdrt(y), O.bdrt(y), f"bdrt {dt} n={n}")

    def test_parity_hash(self, cuda):
        x = np.random.default_rng(0).random((8, 128, 128), dtype=np.float32)
        out = adrt_amd.bdrt(O.adrt(x))
        assert canonical_sha256(out) == golden_hashes()["bdrt_f32_of_adrt_8x128_seed0"]


@gpu
class TestBdrtRows:
    """adrt_hip_bdrt_rows (bdrt for the rows [0, rows) of every quadrant)
    through the C ABI, and the adjoint built on it."""

    @pytest.mark.parametrize("dt,n,b", [("float32", 16, 4), ("float64", 64, 8),
                                        ("float32", 128, 4), ("float64", 256, 2),
                                        ("float32", 512, 3), ("float64", 1024, 1),
                                        ("float32", 2048, 2)])
    def test_rows_equal_bdrt(self, cuda, dt, n, b):
        """Kept rows are bit-identical to the full back-projection's for row
        limits inside the first chunk, at chunk boundaries and at n and 2n-1."""
        import torch
        from adrt_amd import _kernels

        rng = np.random.default_rng(3 * n + b)
        y = torch.from_numpy((rng.standard_normal((b, 4, 2 * n - 1, n)) * 10).astype(dt)).to(cuda)
        full = _kernels.bdrt(y)
        ity = torch.int32 if dt == "float32" else torch.int64
        for rows in sorted({1, min(192, 2 * n - 1), min(384, 2 * n - 1), n - 1 or 1, n,
                            2 * n - 1}):
            got = _kernels.bdrt_rows(y, rows)
            assert torch.equal(got[:, :, :rows].view(ity), full[:, :, :rows].view(ity)), rows

    @pytest.mark.parametrize("dt,n", [("float32", 512), ("float64", 256), ("float32", 2048)])
    def test_adjoint_equals_oracle(self, cuda, dt, n):
        """adjoint(y) = sum over quadrants of truncate(bdrt(y)), equal to the
        oracle's bdrt followed by the same truncation and quadrant sum."""
        import torch
        from adrt_amd import _kernels

        rng = np.random.default_rng(n)
        y = (rng.standard_normal((2, 4, 2 * n - 1, n)) * 10).astype(dt)
        got = _kernels.adjoint(torch.from_numpy(y).to(cuda), 1.0, _kernels.SUM).cpu().numpy()
        t = O.truncate(O.bdrt(y))
        want = ((t[:, 0] + t[:, 1]) + t[:, 2]) + t[:, 3]
        assert np.array_equal(got + 0.0, want + 0.0)


def _delta(size, q, h, s):
    a = np.zeros((4, 2 * size - 1, size))
    a[q, h, s] = 1.0
    return a
