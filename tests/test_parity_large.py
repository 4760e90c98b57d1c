"""Parity at the benchmark sizes: the configurations the bench times.

* The 16-phantom 2048 x 2048 round trip (adrt -> bdrt -> iadrt), float32 and
  float64: every leg byte-equal to the oracle on seeded ellipse phantoms, and
  the exact round trip truncate(iadrt(adrt(x))).mean(-3) == x on phantoms
  quantised to multiples of 1/256 (every partial sum of the transform and of
  its inverse is then exactly representable in either dtype, so the
  reference's inverse returns x exactly: reference tests/test_iadrt.py:185-223
  checks the same property on integer images).  On unquantised data the exact
  inverse amplifies rounding error without bound (the float32 round trip of
  an unquantised 1024 x 1024 phantom is off by about 1e19), which is the
  reference algorithm's own behaviour: there the legs are compared with the
  oracle, not with x.
* iadrt at n = 1024 and 2048 (float32 and float64) and bdrt in float64 at
  1024 and 2048, on standard-normal data, byte-equal to the oracle
  (adrt_cdefs_iadrt.hpp:52-176, adrt_cdefs_bdrt.hpp:55-187).
"""

import numpy as np
import pytest

from helpers import assert_bitwise
from oracle import oracle as O

gpu = pytest.mark.gpu


def _phantoms(cuda, count, n, dt):
    from adrt_amd.phantom import phantoms

    return phantoms(count, n, seed=0, dtype=dt, device=cuda)


@gpu
class TestRoundTrip2048:
    @pytest.mark.parametrize("dt", ["float32", "float64"])
    def test_each_leg_matches_oracle(self, cuda, dt):
        from adrt_amd import _kernels

        print(f"round trip legs 2 x 2048^2 {dt}", flush=True)
        x = _phantoms(cuda, 2, 2048, dt)
        y = _kernels.adrt(x)
        xh = x.cpu().numpy()
        yo = O.adrt(xh)
        assert_bitwise(y.cpu().numpy(), yo, f"adrt {dt} 2048")
        z = _kernels.bdrt(y)
        assert_bitwise(z.cpu().numpy(), O.bdrt(yo), f"bdrt {dt} 2048")
        del z
        w = _kernels.iadrt(y)
        assert_bitwise(w.cpu().numpy(), O.iadrt(yo), f"iadrt {dt} 2048")

    @pytest.mark.parametrize("dt", ["float32", "float64"])
    def test_quantised_round_trip_is_exact(self, cuda, dt):
        import torch

        from adrt_amd import _kernels, utils

        print(f"quantised round trip 2 x 2048^2 {dt}", flush=True)
        x = _phantoms(cuda, 2, 2048, "float64")
        x = (torch.round(x * 256) / 256).to(getattr(torch, dt))
        inv = utils.truncate(_kernels.iadrt(_kernels.adrt(x))).mean(dim=-3)
        assert torch.equal(inv, x)


@gpu
class TestLargeSides:
    @pytest.mark.parametrize("dt", ["float32", "float64"])
    @pytest.mark.parametrize("n", [1024, 2048])
    def test_iadrt_matches_oracle(self, cuda, dt, n):
        import torch

        from adrt_amd import _kernels

        print(f"iadrt n={n} {dt}", flush=True)
        b = 2 if n == 1024 else 1
        y = np.random.default_rng(n + 40).standard_normal((b, 4, 2 * n - 1, n)).astype(dt)
        y.flat[:2] = [-0.0, 0.0]
        got = _kernels.iadrt(torch.from_numpy(y).to(cuda)).cpu().numpy()
        assert_bitwise(got, O.iadrt(y), f"iadrt {dt} n={n}")

    @pytest.mark.parametrize("n", [1024, 2048])
    def test_bdrt_float64_matches_oracle(self, cuda, n):
        import torch

        from adrt_amd import _kernels

        print(f"bdrt n={n} float64", flush=True)
        b = 2 if n == 1024 else 1
        y = np.random.default_rng(n + 41).standard_normal((b, 4, 2 * n - 1, n))
        got = _kernels.bdrt(torch.from_numpy(y).to(cuda)).cpu().numpy()
        assert_bitwise(got, O.bdrt(y), f"bdrt float64 n={n}")
