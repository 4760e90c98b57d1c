"""Autograd through the hand-written kernels (adrt_amd.autograd).

Gradients are checked against central finite differences in float64 and
bit for bit against the reference's adjoint identities restated by the
oracle: A^T g = sum_q truncate(bdrt(g)) (tests/test_bdrt.py:310-325 of the
reference) and bdrt^T = the adrt_step chain (tests/test_bdrt_step.py:76-92).
"""

import numpy as np
import pytest

from helpers import assert_bitwise
from oracle import oracle as O

gpu = pytest.mark.gpu


def test_stage_chain_is_bdrt_adjoint_on_cpu():
    """Pins the identity the bdrt backward uses, on the oracle."""
    rng = np.random.default_rng(0)
    n, L = 16, 4
    g = rng.standard_normal((4, 2 * n - 1, n))
    h = rng.standard_normal((4, 2 * n - 1, n))
    chain = h
    for k in range(L):
        chain = O.adrt_step(chain, k)
    assert abs(np.sum(O.bdrt(g) * h) - np.sum(g * chain)) < 1e-9 * np.sum(np.abs(g * chain))


@pytest.mark.parametrize("shape", [(2, 32, 64), (64, 32), (2, 2, 32, 32), (5,), (2, 24, 24),
                                   (0, 16, 16)])
def test_adrt_rejects_bad_shapes_before_any_device_work(shape):
    """Non-square, non-power-of-two, 1-D, 4-D and empty images are refused on
    the host (a 4-D input used to be read as one image, a non-square one as
    a larger buffer)."""
    import torch
    from adrt_amd import autograd as AG

    with pytest.raises(ValueError):
        AG.adrt(torch.zeros(shape))


@pytest.mark.parametrize("shape", [(4, 30, 16), (3, 31, 16), (2, 4, 31, 15), (2, 2, 4, 31, 16),
                                   (4, 23, 12), (31, 16)])
def test_bdrt_and_lsq_reject_bad_shapes(shape):
    import torch
    from adrt_amd import autograd as AG

    with pytest.raises(ValueError):
        AG.bdrt(torch.zeros(shape))
    with pytest.raises(ValueError):
        AG.lsq_cg(torch.zeros(shape), iters=1)


def test_rejects_host_tensors_and_other_dtypes():
    import torch
    from adrt_amd import autograd as AG

    with pytest.raises(TypeError, match="host tensor"):
        AG.adrt(torch.zeros((2, 16, 16)))
    with pytest.raises(TypeError, match="host tensor"):
        AG.bdrt(torch.zeros((4, 31, 16), dtype=torch.float64))
    with pytest.raises(TypeError, match="torch tensor"):
        AG.adrt(np.zeros((16, 16), dtype=np.float32))


@gpu
class TestAutograd:
    def test_device_dtype_rejected(self, cuda):
        import torch
        from adrt_amd import autograd as AG

        with pytest.raises(TypeError, match="unsupported array dtype"):
            AG.adrt(torch.zeros((16, 16), dtype=torch.float16, device=cuda))
        with pytest.raises(ValueError, match="square"):
            AG.adrt(torch.zeros((2, 16, 32), device=cuda))

    def test_adrt_backward_is_reference_adjoint(self, cuda):
        import torch
        from adrt_amd import autograd as AG

        rng = np.random.default_rng(1)
        x = torch.from_numpy(rng.standard_normal((2, 64, 64))).to(cuda).requires_grad_(True)
        g = rng.standard_normal((2, 4, 127, 64))
        y = AG.adrt(x)
        y.backward(torch.from_numpy(g).to(cuda))
        assert_bitwise(y.detach().cpu().numpy(), O.adrt(x.detach().cpu().numpy()))
        assert_bitwise(x.grad.cpu().numpy(), O.truncate(O.bdrt(g)).sum(axis=-3))

    def test_bdrt_backward_is_stage_chain(self, cuda):
        import torch
        from adrt_amd import autograd as AG

        rng = np.random.default_rng(2)
        y = torch.from_numpy(rng.standard_normal((4, 127, 64))).to(cuda).requires_grad_(True)
        g = rng.standard_normal((4, 127, 64))
        AG.bdrt(y).backward(torch.from_numpy(g).to(cuda))
        want = g
        for k in range(6):
            want = O.adrt_step(want, k)
        assert_bitwise(y.grad.cpu().numpy(), want)

    def test_loss_gradient_finite_differences(self, cuda):
        """0.5 ||adrt(x) - b||^2 at 64 x 64 in float64, central differences
        with eps = 1e-6 along random directions."""
        import torch
        from adrt_amd import autograd as AG

        rng = np.random.default_rng(3)
        b = torch.from_numpy(rng.standard_normal((4, 127, 64))).to(cuda)
        x = torch.from_numpy(rng.standard_normal((64, 64))).to(cuda).requires_grad_(True)

        def loss(z):
            return 0.5 * ((AG.adrt(z) - b) ** 2).sum()

        loss(x).backward()
        eps = 1e-6
        for k in range(4):
            d = torch.from_numpy(rng.standard_normal((64, 64))).to(cuda)
            with torch.no_grad():
                fd = (loss(x + eps * d) - loss(x - eps * d)) / (2 * eps)
            an = (x.grad * d).sum()
            assert abs(float(fd) - float(an)) <= 1e-6 * abs(float(an)), (k, float(fd), float(an))

    def test_gradcheck_small(self, cuda):
        import torch
        from adrt_amd import autograd as AG

        x = torch.rand(1, 8, 8, dtype=torch.float64, device=cuda, requires_grad=True)
        assert torch.autograd.gradcheck(AG.adrt, (x,), eps=1e-6, atol=1e-7)
        y = torch.rand(4, 15, 8, dtype=torch.float64, device=cuda, requires_grad=True)
        assert torch.autograd.gradcheck(AG.bdrt, (y,), eps=1e-6, atol=1e-7)

    def test_lsq_cg_decreases_and_fits(self, cuda):
        import torch
        from adrt_amd import autograd as AG
        from adrt_amd.phantom import phantoms

        img = phantoms(2, 64, seed=4, dtype="float64", device=cuda)
        b = AG.adrt(img).detach()
        x, hist = AG.lsq_cg(b, iters=80)
        losses = torch.stack(hist).cpu().numpy()
        assert np.all(np.diff(losses, axis=0) <= 1e-9 * losses[0])  # CGLS: monotone
        assert np.all(losses[-1] < 1e-12 * losses[0])
        assert float((x - img).abs().max()) < 1e-6

    @pytest.mark.parametrize("dt", ["float32", "float64"])
    def test_lsq_cg_graph_replays_equal_eager(self, cuda, dt):
        """The captured blocks of iterations (replayed 5 times, with the
        host's checks and copies between replays) give exactly the eager
        iterates and loss history."""
        import torch
        from adrt_amd import autograd as AG
        from adrt_amd.phantom import phantoms

        img = phantoms(2, 64, seed=6, dtype=dt, device=cuda)
        b = AG.adrt(img).detach()
        b = b + 0.01 * b.abs().max() * torch.randn(b.shape, dtype=b.dtype, device=cuda,
                                                   generator=torch.Generator(device=cuda).manual_seed(1))
        xg, hg = AG.lsq_cg(b, iters=100, check_every=16, graph=True)
        xe, he = AG.lsq_cg(b, iters=100, check_every=16, graph=False)
        assert len(hg) == len(he) == 101
        assert torch.equal(xg, xe)
        assert all(torch.equal(u, v) for u, v in zip(hg, he))

    def test_lsq_cg_float32_noisy(self, cuda):
        import torch
        from adrt_amd import autograd as AG
        from adrt_amd.phantom import phantoms

        img = phantoms(3, 128, seed=5, dtype="float32", device=cuda)
        b = AG.adrt(img).detach()
        b = b + 0.01 * b.abs().max() * torch.randn_like(b)
        x, hist = AG.lsq_cg(b, iters=40)
        assert x.shape == (3, 128, 128) and x.dtype == torch.float32
        assert float(hist[-1].max()) < float(hist[0].min())
