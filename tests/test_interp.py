"""adrt_amd.utils.interp_to_cart (mirrors the reference's
tests/test_interp_to_cart.py): the extension-level argument checks, and
bit-exact agreement with the reference's interp_adrtcart through the
committed fixtures (tests/golden/, written from the compiled reference) and
the compiled reference itself when it is built."""

import numpy as np
import pytest

import adrt_amd
from helpers import canonical_sha256, golden_hashes, golden_small

gpu = pytest.mark.gpu


class TestInterpToCartCdefs:
    def test_refuses_int32(self):
        with pytest.raises(TypeError, match="int32"):
            adrt_amd._adrt_cdefs.interp_to_cart(np.zeros((4, 15, 8), dtype=np.int32))

    def test_refuses_non_array(self):
        with pytest.raises(TypeError, match="must be a NumPy array or compatible subclass"):
            adrt_amd._adrt_cdefs.interp_to_cart(None)
        with pytest.raises(TypeError, match="must be a NumPy array or compatible subclass"):
            adrt_amd._adrt_cdefs.interp_to_cart(np.arange(4, 7, 4, dtype=np.float32).tolist())

    def test_refuses_fortran_order(self):
        inarr = np.zeros((4, 31, 16), dtype=np.float32, order="F")
        with pytest.raises(ValueError, match=r"must be .*C-order"):
            adrt_amd._adrt_cdefs.interp_to_cart(inarr)

    def test_refuses_c_non_contiguous(self):
        inarr = np.zeros((4, 62, 16), dtype=np.float32, order="F")[:, ::2]
        assert not inarr.flags["C_CONTIGUOUS"]
        with pytest.raises(ValueError, match=r"must be .*contiguous"):
            adrt_amd._adrt_cdefs.interp_to_cart(inarr)

    def test_refuses_byteswapped(self):
        inarr = np.ones((4, 31, 16), dtype=np.float32)
        with pytest.raises(ValueError, match=r"must be .*native byte order"):
            adrt_amd._adrt_cdefs.interp_to_cart(inarr.astype(inarr.dtype.newbyteorder("S")))

    @pytest.mark.parametrize("shape", [(4, 1, 1), (2, 4, 1, 1), (3, 15, 8), (4, 14, 8),
                                       (4, 11, 6), (1, 2, 4, 15, 8)])
    def test_refuses_bad_shapes(self, shape):
        with pytest.raises(ValueError):
            adrt_amd._adrt_cdefs.interp_to_cart(np.zeros(shape, dtype=np.float32))

    def test_wrapper_refuses_non_array(self):
        with pytest.raises(TypeError, match="array must be numpy.ndarray"):
            adrt_amd.utils.interp_to_cart([[1.0]])


@gpu
class TestInterpToCart:
    @pytest.mark.parametrize("dtype", ["float32", "float64"])
    @pytest.mark.parametrize("n", [2, 4, 8, 16])
    def test_matches_golden_vectors(self, cuda, dtype, n):
        g = golden_small()
        key = f"{dtype}_n{n}"
        dat = g[f"{key}_dat"]
        got = adrt_amd.utils.interp_to_cart(dat)
        want = g[f"{key}_interp"]
        assert got.dtype == want.dtype and got.shape == want.shape == (2, n, 4 * n)
        assert got.tobytes() == want.tobytes()
        # unbatched input keeps its rank
        assert adrt_amd.utils.interp_to_cart(dat[0]).tobytes() == want[0].tobytes()

    def test_matches_golden_hashes(self, cuda):
        h = golden_hashes()
        x = np.random.default_rng(0).random((8, 128, 128), dtype=np.float32)
        y = adrt_amd.adrt(x)
        assert canonical_sha256(adrt_amd.utils.interp_to_cart(y)) == h["interp_f32_of_adrt_8x128_seed0"]
        x64 = np.random.default_rng(0).random((2, 256, 256))
        y64 = adrt_amd.adrt(x64)
        assert canonical_sha256(adrt_amd.utils.interp_to_cart(y64)) == h["interp_f64_of_adrt_2x256_seed0"]

    @pytest.mark.parametrize("n", [32, 512, 2048])
    def test_matches_compiled_reference(self, cuda, n):
        from oracle import ref as R

        if not R.available():
            pytest.skip("compiled reference not built (oracle/_ref)")
        a = np.random.default_rng(n).standard_normal((2, 4, 2 * n - 1, n)).astype(np.float32)
        assert adrt_amd.utils.interp_to_cart(a).tobytes() == R.interp_to_cart(a).tobytes()

    def test_device_tensor_stays_on_device(self, cuda):
        import torch

        t = torch.rand((3, 4, 63, 32), device=cuda)
        out = adrt_amd.utils.interp_to_cart(t)
        assert out.is_cuda and out.shape == (3, 32, 128)
        want = adrt_amd.utils.interp_to_cart(t.cpu().numpy())
        assert np.array_equal(out.cpu().numpy(), want)

    def test_zero_outside_and_scaling(self, cuda):
        """Constant data 1: every in-range entry is its scale factor
        sqrt(1 + (si/(n-1))^2) in [1, sqrt(2)], the rest exactly 0."""
        n = 64
        out = adrt_amd.utils.interp_to_cart(np.ones((4, 2 * n - 1, n), dtype=np.float64))
        nz = out[out != 0]
        assert nz.size > 0 and np.all(nz >= 1.0) and np.all(nz <= np.sqrt(2) + 1e-15)


def _cell_centres(a, b, n):
    grid, step = np.linspace(a, b, num=n, endpoint=False, retstep=True)
    return grid + 0.5 * step


def model_interp(adrt_out):
    """interp_to_cart through the host coordinate map (the reference's test
    model): cell-centred theta and t grids, nearest ADRT entry, scale factor,
    zero where the height falls outside the output."""
    n = adrt_out.shape[-1]
    res = np.zeros((*adrt_out.shape[:-3], n, 4 * n))
    ang, off = np.meshgrid(_cell_centres(-0.5 * np.pi, 0.5 * np.pi, 4 * n),
                           _cell_centres(np.sqrt(2) / 2, -np.sqrt(2) / 2, n))
    q, h, s, f = adrt_amd.utils.coord_cart_to_adrt(ang, off, n)
    ok = (h > -1) & (h < 2 * n - 1)
    res[..., ok] = f[ok] * adrt_out[..., q[ok], h[ok], s[ok]]
    return res


@gpu
@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_matches_coordinate_model(cuda, dtype):
    a = np.arange(48, dtype=dtype).reshape((3, 4, 4))
    out = adrt_amd.adrt(a)
    native = adrt_amd.utils.interp_to_cart(out)
    assert native.dtype == a.dtype and native.shape == (3, 4, 16)
    assert np.allclose(model_interp(out), native)
    for i in range(3):
        assert np.array_equal(adrt_amd.utils.interp_to_cart(out[i]), native[i])
