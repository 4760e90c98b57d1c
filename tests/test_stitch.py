"""adrt_amd.utils.stitch_adrt / unstitch_adrt (mirrors the reference's
tests/test_stitch_adrt.py and tests/test_unstitch_adrt.py): seams, quadrant
placement, the zero stencil, removed repeated columns, dtypes, batches,
round trips, and byte equality with a direct NumPy statement of the layout."""

import numpy as np
import pytest

import adrt_amd

gpu = pytest.mark.gpu


def layout_stitch(a, remove_repeated=False):
    """The stitched layout written out directly: quadrant i in column block
    i, odd quadrants flipped in both axes, quadrants 0/1 in the top 2n-1
    rows and 2/3 in the bottom 2n-1 rows, zeros elsewhere."""
    n = a.shape[-1]
    vc = n - 1 if remove_repeated else n
    out = np.zeros((*a.shape[:-3], 3 * n - 2, 4, vc), dtype=a.dtype)
    for i in range(4):
        q = a[..., i, :, :]
        if i % 2:
            q = q[..., ::-1, ::-1]
        q = q[..., :vc]
        rows = slice(0, 2 * n - 1) if i < 2 else slice(n - 1, 3 * n - 2)
        out[..., rows, i, :] = q
    return out.reshape((*a.shape[:-3], 3 * n - 2, 4 * vc))


def check_seams(st):
    n = st.shape[-1] // 4
    for k in (1, 2, 3):
        assert np.array_equal(st[..., :, k * n - 1], st[..., :, k * n])
    assert np.array_equal(st[..., :, -1], st[..., ::-1, 0])


def check_quadrants(st, out):
    n = st.shape[-1] // 4
    assert np.array_equal(st[..., : 2 * n - 1, :n], out[..., 0, :, :])
    assert np.array_equal(st[..., : 2 * n - 1, n : 2 * n], out[..., 1, ::-1, ::-1])
    assert np.array_equal(st[..., -2 * n + 1 :, 2 * n : 3 * n], out[..., 2, :, :])
    assert np.array_equal(st[..., -2 * n + 1 :, 3 * n :], out[..., 3, ::-1, ::-1])


def check_zeros(st):
    n = st.shape[-1] // 4
    assert np.all(st[..., -n + 1 :, : 2 * n] == 0) and np.all(st[..., : n - 1, 2 * n :] == 0)
    ua, ub = np.triu_indices(2 * n - 1, m=n, k=1)
    la, lb = np.tril_indices(2 * n - 1, m=n, k=-n)
    assert np.all(st[..., : 2 * n - 1, :n][..., la, lb] == 0)
    assert np.all(st[..., : 2 * n - 1, n : 2 * n][..., ua, ub] == 0)
    assert np.all(st[..., n - 1 :, 2 * n : 3 * n][..., la, lb] == 0)
    assert np.all(st[..., n - 1 :, 3 * n :][..., ua, ub] == 0)


@pytest.mark.parametrize("shape", [(4, 0, 0), (4, 4, 2), (7, 4)])
def test_stitch_rejects_invalid_sizes(shape):
    with pytest.raises(ValueError, match=r"unsuitable shape .*ADRT output"):
        adrt_amd.utils.stitch_adrt(np.ones(shape, dtype=np.float32))


def test_unstitch_rejects_wrong_dimensions():
    n = 16
    with pytest.raises(ValueError, match=r"unsuitable shape .*ADRT unstitch"):
        adrt_amd.utils.unstitch_adrt(np.zeros((5, 3 * n - 2, 4 * n - 3), dtype=np.float32))
    with pytest.raises(ValueError, match=r"unsuitable shape .*ADRT unstitch"):
        adrt_amd.utils.unstitch_adrt(np.zeros((5, 3 * n - 1, 4 * n), dtype=np.float32))


def test_layout_statement_round_trips():
    """The direct statement used as the reference layout is consistent with
    itself (CPU only): quadrant blocks and seams of a real ADRT output."""
    from oracle import oracle as O

    out = O.adrt(np.arange(64, dtype=np.float32).reshape(8, 8))
    st = layout_stitch(out)
    check_seams(st)
    check_quadrants(st, out)
    check_zeros(st)


@gpu
class TestStitch:
    @pytest.mark.parametrize("batch", [(), (3,), (2, 3)])
    def test_adrt_output(self, cuda, batch):
        n = 16
        img = np.arange(int(np.prod(batch, dtype=int)) * n * n, dtype=np.float32)
        out = adrt_amd.adrt(img.reshape((-1, n, n))).reshape((*batch, 4, 2 * n - 1, n))
        st = adrt_amd.utils.stitch_adrt(out)
        assert st.shape == (*batch, 3 * n - 2, 4 * n)
        check_seams(st)
        check_quadrants(st, out)
        check_zeros(st)
        assert st.tobytes() == layout_stitch(out).tobytes()

    @pytest.mark.parametrize("batch", [(), (3,), (2, 3)])
    def test_remove_repeated_drops_the_seam_columns(self, cuda, batch):
        n = 8
        img = np.arange(int(np.prod(batch, dtype=int)) * n * n, dtype=np.float32)
        out = adrt_amd.adrt(img.reshape((-1, n, n))).reshape((*batch, 4, 2 * n - 1, n))
        st = adrt_amd.utils.stitch_adrt(out, remove_repeated=True)
        assert st.shape == (*batch, 3 * n - 2, 4 * n - 4)
        full = np.delete(adrt_amd.utils.stitch_adrt(out), [i * n - 1 for i in range(1, 5)], axis=-1)
        assert np.array_equal(st, full)
        assert st.tobytes() == layout_stitch(out, True).tobytes()

    @pytest.mark.parametrize("dtype", ["float32", "float64", "int32", "int64", "int16", "uint8"])
    @pytest.mark.parametrize("n", [1, 3, 8])
    def test_dtypes_and_any_size(self, cuda, dtype, n):
        a = (np.random.default_rng(n).random((4, 2 * n - 1, n)) * 100).astype(dtype)
        st = adrt_amd.utils.stitch_adrt(a)
        assert st.dtype == a.dtype and st.shape == (3 * n - 2, 4 * n)
        assert st.tobytes() == layout_stitch(a).tobytes()

    @pytest.mark.parametrize("dtype", ["float32", "float64", "int32", "int64"])
    @pytest.mark.parametrize("remove_repeated", [True, False])
    def test_unstitch_round_trip(self, cuda, dtype, remove_repeated):
        n = 16
        out = adrt_amd.adrt(np.arange(n * n, dtype=np.float64).reshape(n, n)).astype(dtype)
        st = adrt_amd.utils.stitch_adrt(out, remove_repeated=remove_repeated)
        back = adrt_amd.utils.unstitch_adrt(st)
        assert back.dtype == out.dtype and back.shape == out.shape
        assert np.array_equal(back, out)

    @pytest.mark.parametrize("remove_repeated", [True, False])
    def test_unstitch_round_trip_batch(self, cuda, remove_repeated):
        n = 16
        out = adrt_amd.adrt(np.arange(6 * n * n, dtype=np.float32).reshape(6, n, n))
        out = out.reshape((2, 3, *out.shape[1:]))
        back = adrt_amd.utils.unstitch_adrt(adrt_amd.utils.stitch_adrt(out, remove_repeated=remove_repeated))
        assert back.shape == out.shape and np.array_equal(back, out)

    def test_device_tensors(self, cuda):
        import torch

        a = torch.rand((2, 4, 31, 16), device=cuda)
        st = adrt_amd.utils.stitch_adrt(a, remove_repeated=True)
        assert st.is_cuda and st.shape == (2, 46, 60)
        assert np.array_equal(st.cpu().numpy(), layout_stitch(a.cpu().numpy(), True))
