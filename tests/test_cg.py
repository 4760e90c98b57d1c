"""Device CG inverse (adrt_amd.cg), the reference's documented recipe
(docs/examples.cginverse.md:40-67, docs/examples.tomography.md:63-71).

The operator is checked bit for bit against the oracle's restatement of the
recipe's operator; the solver, which the reference has no test for, is
checked against the CPU float64 recipe to a tolerance, against its own
stopping rule, and for batch/unbatched agreement."""

import numpy as np
import pytest

from adrt_amd import cg
from helpers import assert_bitwise
from oracle import oracle as O

gpu = pytest.mark.gpu


def recipe_image(n=16):
    """The input image of the reference recipe (docs/examples.cginverse.md:73-81)."""
    xs = np.linspace(-1, 1, n)
    x, y = np.meshgrid(xs, xs)
    img = 0.5 * ((np.abs(x - 0.25) + np.abs(y)) < 0.7).astype(np.float32)
    img[:, 3] = 1
    img[1, :] = 1
    return img


def noisy_data(n=16, seed=0, scale=1e-4, dtype=np.float64):
    a = O.adrt(recipe_image(n).astype(dtype))
    return (a + np.random.default_rng(seed).normal(scale=scale, size=a.shape)).astype(dtype)


def _scipy_recipe(b, rtol=1e-5, x0=None):
    """The recipe's own solver (docs/examples.cginverse.md:58-67):
    scipy.sparse.linalg.cg on a LinearOperator over the oracle's normal
    operator, at SciPy's default rtol; returns (x, info, iterations)."""
    from scipy.sparse.linalg import LinearOperator, cg as scipy_cg

    n = b.shape[-1]
    op = LinearOperator((n * n, n * n), dtype=np.float64,
                        matvec=lambda v: O.normal_op(v.reshape(n, n)).ravel())
    rhs = O.truncate(O.bdrt(b)).mean(axis=0).ravel()
    count = [0]

    def cb(_):
        count[0] += 1

    x, info = scipy_cg(op, rhs, x0=x0, rtol=rtol, callback=cb)
    return x.reshape(n, n), info, count[0]


@pytest.mark.parametrize("n,seed", [(16, 0), (16, 1), (32, 2)])
def test_oracle_cg_matches_scipy_cg(n, seed):
    """The oracle's float64 CG (the checker of the device solver) against
    scipy.sparse.linalg.cg, the solver the recipe calls, on the recipe's
    operator: the same iteration count and the same solution."""
    b = noisy_data(n, seed=seed, scale=1e-2)
    xs, info, its = _scipy_recipe(b)
    xo, ko, hist = O.cg_normal(b, rtol=1e-5, maxiter=10000)
    assert info == 0 and its > 3
    assert ko == its, (ko, its)
    assert hist[-1] <= 1e-5
    assert np.max(np.abs(xo - xs)) <= 1e-9 * np.max(np.abs(xs))


def test_refuses_batch():
    with pytest.raises(ValueError, match="batch dimension not supported for iadrt_cg"):
        cg.iadrt_cg(np.zeros((2, 4, 31, 16)))


@gpu
class TestCG:
    @pytest.mark.parametrize("dt", ["float32", "float64"])
    def test_normal_operator_matches_recipe(self, cuda, dt):
        import torch

        x = np.random.default_rng(3).standard_normal((3, 32, 32)).astype(dt)
        got = cg.ADRTNormalOperator(32)(torch.from_numpy(x).to(cuda)).cpu().numpy()
        assert_bitwise(got, O.normal_op(x))

    def test_ridge_operator(self, cuda):
        import torch

        x = torch.from_numpy(np.random.default_rng(4).standard_normal((2, 16, 16))).to(cuda)
        m = cg.ADRTNormalOperator(16)(x)
        r = cg.ADRTRidgeOperator(16, ridge_param=2.5)(x)
        assert torch.equal(r, m + 2.5 * x)

    def test_operator_symmetric(self, cuda):
        import torch

        rng = np.random.default_rng(5)
        x = torch.from_numpy(rng.standard_normal((64, 64))).to(cuda)
        y = torch.from_numpy(rng.standard_normal((64, 64))).to(cuda)
        op = cg.ADRTNormalOperator(64)
        lhs = float((op(x) * y).sum())
        rhs = float((x * op(y)).sum())
        assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), 1.0)

    @pytest.mark.parametrize("n,seed", [(16, 0), (32, 2)])
    def test_matches_scipy_cg(self, cuda, n, seed):
        """The device solver against scipy.sparse.linalg.cg on the recipe's
        operator at SciPy's default rtol (docs/examples.cginverse.md:58-67):
        the same number of operator applications, a true residual (the
        oracle's float64 operator) below rtol, and the same solution to the
        tolerance two converged runs can differ by.  The device's float64
        inner products sum in another order than NumPy's, and CG amplifies
        that rounding over its iterations: measured 6e-7 of max|x| at n = 16,
        where two runs with identical arithmetic agree to 1e-9 (the CPU test
        above)."""
        import torch

        b = noisy_data(n, seed=seed, scale=1e-2)
        xs, info, its = _scipy_recipe(b)
        op = cg.ADRTNormalOperator(n)
        rhs = cg.normal_rhs(torch.from_numpy(b).to(cuda))
        x, dinfo, iters = cg.cg_solve(op, rhs)
        assert info == 0 and int(dinfo) == 0
        assert iters == its, (iters, its)
        xd = x.cpu().numpy()
        rhs_ref = O.truncate(O.bdrt(b)).mean(axis=0)
        res = np.linalg.norm(rhs_ref - O.normal_op(xd)) / np.linalg.norm(rhs_ref)
        assert res < 1e-5, res
        assert np.max(np.abs(xd - xs)) <= 1e-5 * np.max(np.abs(xs))

    def test_matches_cpu_recipe(self, cuda):
        b = noisy_data(16)
        x = cg.iadrt_cg(b, rtol=1e-8)
        want, _, _ = O.cg_normal(b, rtol=1e-8, maxiter=10000)
        assert x.shape == (16, 16) and x.dtype == np.float64
        assert np.max(np.abs(x - want)) <= 1e-5 * np.max(np.abs(want))

    @pytest.mark.parametrize("dt", ["float32", "float64"])
    def test_stopping_rule(self, cuda, dt):
        import torch

        b = noisy_data(32, dtype=dt)
        tb = torch.from_numpy(b).to(cuda)
        op = cg.ADRTNormalOperator(32)
        rhs = cg.normal_rhs(tb)
        x, info, iters = cg.cg_solve(op, rhs, rtol=1e-4)
        assert int(info) == 0 and iters > 0
        res = (rhs.double() - op(x).double()).norm() / rhs.double().norm()
        assert float(res) < 1e-4 * 1.01

    def test_recovers_image(self, cuda):
        img = recipe_image(16).astype(np.float64)
        x = cg.iadrt_cg(O.adrt(img), rtol=1e-10)
        assert np.max(np.abs(x - img)) < 1e-6

    def test_batch_matches_single(self, cuda):
        import torch

        bs = np.stack([noisy_data(32, seed=s) for s in range(3)])
        tb = torch.from_numpy(bs).to(cuda)
        op = cg.ADRTNormalOperator(32)
        rhs = cg.normal_rhs(tb)
        xb, infob, _ = cg.cg_solve(op, rhs, rtol=1e-6)
        for i in range(3):
            xi, infoi, _ = cg.cg_solve(op, rhs[i], rtol=1e-6)
            assert int(infob[i]) == 0 and int(infoi) == 0
            # per-image scalars: the same iterates up to the float64 sum order
            assert torch.allclose(xb[i], xi, rtol=1e-9, atol=1e-12)

    def test_zero_rhs(self, cuda):
        x = cg.iadrt_cg(np.zeros((4, 31, 16)))
        assert np.all(x == 0)

    def test_not_converged_raises(self, cuda):
        with pytest.raises(ValueError, match=r"convergence failed \(cg status 1\)"):
            cg.iadrt_cg(noisy_data(16), maxiter=1)

    def test_device_tensor_stays_on_device(self, cuda):
        import torch

        tb = torch.from_numpy(noisy_data(16)).to(cuda)
        x = cg.iadrt_cg(tb, rtol=1e-6)
        assert x.is_cuda and x.shape == (16, 16)
