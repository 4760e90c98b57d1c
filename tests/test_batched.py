"""Uneven-size batches (adrt_amd.batched): autotuned kernel variants and the
forward + back-projection pipeline captured as one hipGraph.

Every configuration the tuner can select must give the oracle's results bit
for bit (after canonicalising the sign of zero), and a graph replay must give
exactly what the eager calls give, on every replay."""

import numpy as np
import pytest

from adrt_amd import _lib, batched
from helpers import assert_bitwise
from oracle import oracle as O

gpu = pytest.mark.gpu
MIB = 1 << 20


def test_candidates_streaming_only_outside_fused_range():
    assert batched.candidates(_lib.OP_ADRT, 32, 100, 1) == [(_lib.VARIANT_STREAMING, 0)]
    assert batched.candidates(_lib.OP_BDRT, 8192, 2, 1) == [(_lib.VARIANT_STREAMING, 0)]


def test_candidates_alt_split_for_odd_log2():
    per = 4 * MIB
    odd = batched.candidates(_lib.OP_ADRT, 128, 1000, per)
    even = batched.candidates(_lib.OP_ADRT, 256, 1000, per)
    assert {v for v, _ in odd} == {0, 1, 2}
    assert {v for v, _ in even} == {0, 2}


def test_candidates_budgets_distinct_groups():
    per = 40 * MIB
    # 16 MiB and 64 MiB both give one image per group at 40 MiB per image
    got = [b for v, b in batched.candidates(_lib.OP_BDRT, 1024, 4, per) if v == 0]
    groups = [min(max(b // per, 1), 4) for b in got]
    assert len(groups) == len(set(groups)) and 1 in groups and 4 in groups


@pytest.fixture
def reset_config():
    yield
    for op in (_lib.OP_ADRT, _lib.OP_BDRT, _lib.OP_IADRT):
        for L in range(0, 13):
            batched.set_config(op, 1 << L, _lib.VARIANT_DEFAULT, 0)


@gpu
class TestConfig:
    def test_set_get_roundtrip(self, cuda, reset_config):
        batched.set_config(_lib.OP_ADRT, 128, _lib.VARIANT_ALT_SPLIT, 64 * MIB)
        assert batched.get_config(_lib.OP_ADRT, 128) == (_lib.VARIANT_ALT_SPLIT, 64 * MIB)
        batched.set_config(_lib.OP_ADRT, 128, _lib.VARIANT_DEFAULT, 0)
        assert batched.get_config(_lib.OP_ADRT, 128) == (_lib.VARIANT_DEFAULT, 136 * MIB)

    def test_iadrt_variants_roundtrip(self, cuda, reset_config):
        """The inverse's fused-tail variants are set per size like the
        forward's (the group budget does not apply to it)."""
        for v in (_lib.VARIANT_STREAMING, _lib.VARIANT_ALT_SPLIT, _lib.VARIANT_DEFAULT):
            batched.set_config(_lib.OP_IADRT, 32, v, 0)
            assert batched.get_config(_lib.OP_IADRT, 32)[0] == v

    def test_invalid_config(self, cuda, reset_config):
        lib = _lib.load()
        assert lib.adrt_hip_set_config(_lib.OP_ADRT, 32, _lib.VARIANT_ALT_SPLIT, 0) == _lib.ERR_ARG
        assert lib.adrt_hip_set_config(7, 64, 0, 0) == _lib.ERR_ARG
        assert lib.adrt_hip_set_config(_lib.OP_IADRT, 48, 0, 0) == _lib.ERR_ARG
        assert lib.adrt_hip_set_config(_lib.OP_ADRT, 96, 0, 0) == _lib.ERR_ARG
        assert lib.adrt_hip_set_config(_lib.OP_ADRT, 64, 3, 0) == _lib.ERR_ARG

    def test_workspace_follows_config(self, cuda, reset_config):
        import torch

        lib = _lib.load()
        small = lib.adrt_hip_workspace_bytes(_lib.OP_ADRT, _lib.F32, 8, 256)
        batched.set_config(_lib.OP_ADRT, 256, _lib.VARIANT_STREAMING, 0)
        stream_ws = lib.adrt_hip_workspace_bytes(_lib.OP_ADRT, _lib.F32, 8, 256)
        assert stream_ws != small
        # a workspace sized for the old configuration is refused, not overrun
        x = torch.rand((8, 256, 256), device=cuda)
        out = torch.empty((8, 4, 511, 256), device=cuda)
        ws = torch.empty(min(small, stream_ws) - 1, dtype=torch.uint8, device=cuda)
        st = lib.adrt_hip_adrt(x.data_ptr(), out.data_ptr(), ws.data_ptr(), ws.numel(), _lib.F32,
                               8, 256, torch.cuda.current_stream().cuda_stream)
        assert st == _lib.ERR_WORKSPACE

    @pytest.mark.parametrize("n", [64, 128, 512, 2048])
    def test_every_variant_matches_oracle(self, cuda, reset_config, n):
        import torch
        from adrt_amd import _kernels

        b = 3 if n < 2048 else 1
        x = np.random.default_rng(n).standard_normal((b, n, n)).astype(np.float32)
        want = O.adrt(x)
        back_want = O.bdrt(want)
        xt = torch.from_numpy(x).to(cuda)
        per = max(int(_lib.load().adrt_hip_workspace_bytes(_lib.OP_ADRT, _lib.F32, 1, n)), 1)
        for variant, budget in batched.candidates(_lib.OP_ADRT, n, b, per):
            batched.set_config(_lib.OP_ADRT, n, variant, budget)
            batched.set_config(_lib.OP_BDRT, n, variant, budget)
            got = _kernels.adrt(xt)
            assert_bitwise(got.cpu().numpy(), want, f"adrt n={n} variant={variant} budget={budget}")
            back = _kernels.bdrt(got)
            assert_bitwise(back.cpu().numpy(), back_want, f"bdrt n={n} variant={variant}")


@gpu
class TestUnevenBatch:
    SIZES = [32, 64, 64, 128, 256, 32, 512, 128, 32, 64, 1024, 16]

    def _images(self, cuda):
        import torch

        g = torch.Generator(device=cuda)
        g.manual_seed(0)
        return [torch.rand((n, n), generator=g, device=cuda) for n in self.SIZES]

    def test_graph_replays_match_eager_and_oracle(self, cuda, reset_config):
        import torch
        from adrt_amd import _kernels

        plan = batched.UnevenBatch(self.SIZES, reps=1)
        imgs = self._images(cuda)
        plan.load(imgs)
        plan.capture()
        eager_f = [_kernels.adrt(im.contiguous()) for im in imgs]
        eager_b = [_kernels.bdrt(f) for f in eager_f]
        for rep in range(3):
            for n in plan.classes:
                plan.fwd[n].fill_(np.nan)
                plan.back[n].fill_(np.nan)
            plan.replay()
            # eager work between replays
            torch.cuda.synchronize()
            _ = _kernels.adrt(imgs[0].contiguous())
            for i in range(len(self.SIZES)):
                assert torch.equal(plan.forward(i), eager_f[i]), (rep, i)
                assert torch.equal(plan.backprojection(i), eager_b[i]), (rep, i)
        for i in (0, 6, 10):
            x = imgs[i].cpu().numpy()
            assert_bitwise(plan.forward(i).cpu().numpy(), O.adrt(x))
            assert_bitwise(plan.backprojection(i).cpu().numpy(), O.bdrt(O.adrt(x)))

    def test_new_images_through_same_graph(self, cuda, reset_config):
        import torch
        from adrt_amd import _kernels

        plan = batched.UnevenBatch(self.SIZES, autotune=False)
        plan.load(self._images(cuda))
        plan.capture()
        imgs = [torch.full((n, n), 0.5, device=cuda) + i for i, n in enumerate(self.SIZES)]
        plan.load(imgs)
        plan.replay()
        torch.cuda.synchronize()
        for i, im in enumerate(imgs):
            assert torch.equal(plan.forward(i), _kernels.adrt(im))

    def test_rejects_bad_input(self, cuda):
        with pytest.raises(ValueError, match="power of two"):
            batched.UnevenBatch([32, 48], autotune=False)
        plan = batched.UnevenBatch([32, 64], autotune=False)
        with pytest.raises(ValueError, match="expected 2 images"):
            plan.load([])
        with pytest.raises(RuntimeError, match="capture"):
            plan.replay()


def test_lane_assignment_longest_first():
    """Size classes go to at most 4 streams (the hardware queues), costliest
    first onto the least-loaded stream; every class exactly once."""
    plan = batched.UnevenBatch.__new__(batched.UnevenBatch)
    plan.classes = [32, 64, 128, 256, 512, 1024]
    plan.count = {32: 479, 64: 267, 128: 142, 256: 65, 512: 32, 1024: 15}
    ms = {32: (0.25, 0.25), 64: (0.06, 0.1), 128: (0.1, 0.15), 256: (0.15, 0.3),
          512: (0.3, 0.5), 1024: (0.5, 1.0)}
    plan.tuning = [{"op": op, "n": n, "choice": {"ms": ms[n][i]}}
                   for n in plan.classes for i, op in enumerate(("adrt", "bdrt"))]
    lanes = plan._assign_lanes()
    assert len(lanes) == 4
    assert sorted(n for lane in lanes for n in lane) == plan.classes
    assert [1024] in lanes                      # the costliest class runs alone
    costs = [sum(sum(ms[n]) for n in lane) for lane in lanes]
    assert max(costs) == sum(ms[1024])
    for lane in lanes:                          # costliest first within a stream
        assert [sum(ms[n]) for n in lane] == sorted((sum(ms[n]) for n in lane), reverse=True)
    # untuned: bytes plus launch latency, same properties
    plan.tuning = []
    lanes = plan._assign_lanes()
    assert sorted(n for lane in lanes for n in lane) == plan.classes and len(lanes) == 4
