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
        full-batch solve's iterates and iteration counts exactly: an image's
        float64 dot products do not depend on the batch it is in."""
        import torch
        from adrt_amd import learned as LN

        solver = LN.LearnedSolver(n=64, device=cuda)
        solver.fit_preconditioner(count=4)
        x, b = solver.data(seed=8, start=0, count=24, noise_seed=9)
        for prec in (None, solver.pyramid):
            xf, itf, cf, af = LN.pcg(LN.normal_op, b, prec=prec, rtol=1e-5, compact=0)
            xc, itc, cc, ac = LN.pcg(LN.normal_op, b, prec=prec, rtol=1e-5, compact=0.9)
            assert bool(cf.all()) and bool(cc.all())
            assert torch.equal(itf, itc)
            assert torch.equal(xf, xc)
            assert ac <= af


def host_scalar_pcg(op, b, prec, rtol, maxiter):
    """The batched PCG iteration with its per-image scalar logic in torch
    operations between the device vector kernels, one host check per
    iteration: the form the device-side pcg_update / pcg_direction replace.
    Same float64 scalars, same rounding of alpha and beta to the element
    type, same breakdown rule."""
    import torch
    from adrt_amd import _kernels

    bd = _kernels.batched_dot
    x = torch.zeros_like(b)
    r = b.clone()
    tol = rtol * bd(b, b).sqrt()
    rr = bd(r, r)
    use_p = torch.full((b.shape[0],), prec is not None, dtype=torch.bool, device=b.device)
    if prec is None:
        z, rz = r, rr.clone()
    else:
        z = prec(r)
        rz = bd(r, z)
        bad = rz <= 0
        use_p &= ~bad
        z = torch.where(bad[:, None, None], r, z)
        rz = torch.where(bad, rr, rz)
    p = z.clone()
    active = rr.sqrt() > tol
    iters = torch.zeros(b.shape[0], dtype=torch.int64, device=b.device)
    one = torch.ones_like(rz)
    zero = torch.zeros_like(rz)
    for _ in range(maxiter):
        if not bool(active.any()):
            break
        q = op(p)
        pq = bd(p, q)
        alpha = torch.where(active & (pq > 0), rz / torch.where(pq > 0, pq, one), zero)
        rr_new = _kernels.cg_update(x, r, p, q, alpha)
        iters += active.long()
        new_active = active & (rr_new.sqrt() > tol)
        if prec is None:
            zz, rz_new = r, rr_new
            beta = rz_new / torch.where(rz != 0, rz, one)
        else:
            zz = torch.where(use_p[:, None, None], prec(r), r)
            rz_new = bd(r, zz)
            bad = new_active & use_p & ((rz_new <= 0) | (pq <= 0))
            use_p &= ~bad
            zz = torch.where(bad[:, None, None], r, zz)
            rz_new = torch.where(bad, rr_new, rz_new)
            beta = torch.where(bad, zero, rz_new / torch.where(rz != 0, rz, one))
        _kernels.cg_direction(p, zz, beta)
        rz, active = rz_new, new_active
    return x, iters, ~active


class TestPcgStateLayout:
    def test_views_tile_the_buffer(self):
        """The state's fields are disjoint views that tile the library's
        state size in the documented order (no device needed)."""
        import torch
        from adrt_amd import _kernels, _lib

        for B in (1, 3, 1000):
            st = _kernels.PcgState(B, torch.device("cpu"))
            assert st.buf.numel() == _lib.load().adrt_hip_pcg_state_bytes(B)
            spans = []
            for k in _kernels.PcgState.F64 + _kernels.PcgState.I32 + ("count",):
                v = getattr(st, k)
                off = v.data_ptr() - st.buf.data_ptr()
                spans.append((off, off + v.numel() * v.element_size()))
                assert v.numel() == (1 if k == "count" else B)
            assert spans[0][0] == 0
            for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
                assert a1 == b0
            assert spans[-1][1] <= st.buf.numel()


@gpu
class TestDevicePCG:
    def test_update_direction_kernels(self, cuda):
        """pcg_update / pcg_direction on a hand-made state: active and
        inactive images, a preconditioned breakdown (r.z < 0), an image on
        plain CG and an image with pq <= 0, against the same arithmetic in
        float64 torch operations."""
        import torch
        from adrt_amd import _kernels

        rng = np.random.default_rng(11)
        B, n = 6, 32
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
        x, r, p = (t(rng.standard_normal((B, n, n)).astype(np.float32)) for _ in range(3))
        q = p * 2.0 + t(rng.standard_normal((B, n, n)).astype(np.float32)) * 0.1
        q[4] = -p[4]                          # pq < 0
        z = r * 0.5
        z[1] = -r[1]                          # r.z < 0: breakdown
        st = _kernels.PcgState(B, cuda)
        st.rz.copy_(t(rng.uniform(1.0, 2.0, B)))
        st.rz[5] = 0.0                        # rz == 0 reads as 1
        st.tol.copy_(t(np.array([1e-3, 1e-3, 1e9, 1e-3, 1e-3, 1e-3])))  # image 2 converges
        st.active.copy_(t(np.array([1, 1, 1, 0, 1, 1], np.int32)))      # image 3 inactive
        st.use_p.copy_(t(np.array([1, 1, 1, 1, 1, 0], np.int32)))       # image 5 plain
        st.iters.copy_(t(np.array([4, 4, 4, 9, 4, 4], np.int32)))
        x0, r0, p0 = x.clone(), r.clone(), p.clone()
        rz0 = st.rz.clone()
        act0 = st.active.bool().clone()
        ws = _kernels.cg_workspace(x)
        _kernels.pcg_update(x, r, p, q, st, ws)
        pq = _kernels.batched_dot(p0, q)
        alpha = torch.where(act0 & (pq > 0), rz0 / torch.where(pq > 0, pq, 1.0), 0.0)
        a = alpha.float()[:, None, None]
        assert torch.equal(x, x0 + a * p0) and torch.equal(r, r0 - a * q)
        rr = _kernels.batched_dot(r, r)
        assert torch.equal(st.rr, rr) and torch.equal(st.pq, pq)
        nxt = act0 & (rr.sqrt() > st.tol)
        assert st.next.tolist() == nxt.int().tolist()
        assert st.iters.tolist() == [5, 5, 5, 9, 5, 5]
        use0 = st.use_p.bool().clone()
        _kernels.pcg_direction(p, r, z, st, ws)
        rzd = _kernels.batched_dot(r, z)
        bad = nxt & use0 & ((rzd <= 0) | (pq <= 0))
        use1 = use0 & ~bad
        rzn = torch.where(use1, rzd, rr)
        denom = torch.where(rz0 != 0, rz0, 1.0)
        beta = torch.where(bad, 0.0, rzn / denom)
        assert bool(bad[1]) and bool(bad[4]) and not bool(bad[0])
        assert st.use_p.tolist() == use1.int().tolist()
        assert torch.equal(st.rz, rzn) and torch.equal(st.beta, beta)
        sel = torch.where(use1[:, None, None], z, r)
        assert torch.equal(p, sel + beta.float()[:, None, None] * p0)
        assert st.active.tolist() == nxt.int().tolist()
        assert int(st.count) == int(nxt.sum())

    def test_matches_host_scalar_iteration(self, cuda):
        """The device-scalar solver gives the host-scalar iteration's
        iterates bit for bit: plain CG, the pyramid preconditioner, and a
        preconditioner that turns indefinite on two images at its third call
        (the restart as plain CG)."""
        import torch
        from adrt_amd import learned as LN

        solver = LN.LearnedSolver(n=64, device=cuda)
        solver.fit_preconditioner(count=4)
        _, b = solver.data(seed=5, start=0, count=6, noise_seed=6)

        def flipping():
            calls = [0]
            sign = torch.ones(6, 1, 1, device=cuda)
            sign[1] = sign[4] = -1.0

            def f(r):
                calls[0] += 1
                out = solver.pyramid(r)
                return out * sign if calls[0] == 3 else out
            return f

        for name in ("plain", "pyramid", "flip"):
            prec = {"plain": None, "pyramid": solver.pyramid}.get(name)
            pa = flipping() if name == "flip" else prec
            pb = flipping() if name == "flip" else prec
            xr, ir, cr = host_scalar_pcg(LN.normal_op, b, pa, 1e-5, 300)
            xd, idv, cd, _ = LN.pcg(LN.normal_op, b, prec=pb, rtol=1e-5, maxiter=300, compact=0)
            assert bool(cr.all()), name
            assert torch.equal(ir, idv) and torch.equal(cr, cd), name
            assert torch.equal(xr, xd), name
