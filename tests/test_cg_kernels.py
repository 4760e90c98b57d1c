"""Device kernels of the solvers: the fused truncate + quadrant combine, the
batched float64 dot products and CG updates, and the graph-captured PCG
iteration (replays must reproduce the eager iteration exactly)."""

import numpy as np
import pytest

from oracle import oracle as O

gpu = pytest.mark.gpu


def np_truncate_combine(a, divisor, mode):
    """mean (mode 0) or sum (mode 1) over quadrants of truncate(a) / divisor,
    with the sum order ((q0 + q1) + q2) + q3 (core.py:329 for the mean)."""
    t = O.truncate(a)
    if divisor != 1.0:
        t = t / t.dtype.type(divisor)
    acc = ((t[..., 0, :, :] + t[..., 1, :, :]) + t[..., 2, :, :]) + t[..., 3, :, :]
    return acc / acc.dtype.type(4) if mode == 0 else acc


@gpu
class TestTruncateCombine:
    @pytest.mark.parametrize("n", [1, 2, 16, 32, 64, 256])
    @pytest.mark.parametrize("dt", ["float32", "float64"])
    @pytest.mark.parametrize("mode,divisor", [(0, 1.0), (1, 1.0), (0, 3.0)])
    def test_bitwise(self, cuda, n, dt, mode, divisor):
        import torch
        from adrt_amd import _kernels

        a = np.random.default_rng(n).standard_normal((3, 4, 2 * n - 1, n)).astype(dt)
        got = _kernels.truncate_combine(torch.from_numpy(a).to(cuda), divisor, mode)
        want = np_truncate_combine(a, divisor, mode)
        assert got.cpu().numpy().tobytes() == want.tobytes()


@gpu
class TestCGKernels:
    @pytest.mark.parametrize("shape", [(3, 16, 16), (1000, 8, 8), (2, 512, 512), (4, 70001)])
    def test_dot(self, cuda, shape):
        import torch
        from adrt_amd import _kernels

        rng = np.random.default_rng(len(shape))
        u = rng.standard_normal(shape).astype(np.float32)
        v = rng.standard_normal(shape).astype(np.float32)
        got = _kernels.batched_dot(torch.from_numpy(u).to(cuda), torch.from_numpy(v).to(cuda))
        want = (u.astype(np.float64) * v.astype(np.float64)).reshape(shape[0], -1).sum(axis=1)
        assert np.allclose(got.cpu().numpy(), want, rtol=1e-12, atol=1e-12)

    @pytest.mark.parametrize("dt", ["float32", "float64"])
    def test_update_and_direction_match_unfused(self, cuda, dt):
        import torch
        from adrt_amd import _kernels

        rng = np.random.default_rng(3)
        B = 5
        x, r, p, q, z = (torch.from_numpy(rng.standard_normal((B, 64, 64)).astype(dt)).to(cuda)
                         for _ in range(5))
        alpha = torch.from_numpy(rng.standard_normal(B)).to(cuda)
        beta = torch.from_numpy(rng.standard_normal(B)).to(cuda)
        a = alpha.to(x.dtype)[:, None, None]
        c = beta.to(x.dtype)[:, None, None]
        want_x, want_r = x + a * p, r - a * q
        want_p = z + c * p
        rr = _kernels.cg_update(x, r, p, q, alpha)
        assert torch.equal(x, want_x) and torch.equal(r, want_r)
        want_rr = (want_r.double() ** 2).sum(dim=(1, 2))
        assert torch.allclose(rr, want_rr, rtol=1e-12)
        _kernels.cg_direction(p, z, beta)
        assert torch.equal(p, want_p)


@gpu
class TestGraphPCG:
    def test_graph_replays_equal_eager(self, cuda):
        """The captured iteration, replayed with host checks in between,
        gives exactly the eager iterates (plain and preconditioned)."""
        import torch
        from adrt_amd import learned as LN

        solver = LN.LearnedSolver(n=64, device=cuda)
        solver.fit_preconditioner(count=4)
        _, b = solver.data(seed=5, start=0, count=4, noise_seed=6)
        for prec in (None, solver.pyramid):
            xe, ie, ce, ae = LN.pcg(LN.normal_op, b, prec=prec, rtol=1e-5, graph=False)
            torch.cuda.synchronize()
            xg, ig, cg_, ag = LN.pcg(LN.normal_op, b, prec=prec, rtol=1e-5, graph=True)
            assert ae == ag and torch.equal(ie, ig) and torch.equal(ce, cg_)
            assert torch.equal(xe, xg)

    def test_compaction_matches_full_batch(self, cuda):
        """Gathering the still-running images into smaller batches gives the
        full-batch solve's iterates (up to the float64 sum order of the
        per-image dot products, which depends on the batch shape)."""
        import torch
        from adrt_amd import learned as LN

        solver = LN.LearnedSolver(n=64, device=cuda)
        solver.fit_preconditioner(count=4)
        x, b = solver.data(seed=8, start=0, count=24, noise_seed=9)
        for prec in (None, solver.pyramid):
            xf, itf, cf, af = LN.pcg(LN.normal_op, b, prec=prec, rtol=1e-5, compact=0)
            xc, itc, cc, ac = LN.pcg(LN.normal_op, b, prec=prec, rtol=1e-5, compact=0.9)
            assert bool(cf.all()) and bool(cc.all())
            assert int((itf - itc).abs().max()) <= 1
            assert float((xf - xc).norm() / xf.norm()) < 1e-5
            assert ac <= af
