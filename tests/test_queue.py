"""The persistent stage-ordered queue (adrt_hip_queue_*, batched.QueuePlan):
forward ADRT + back-projection of stacks of many power-of-two sides n <= 1024
in one launch.

Every element comes from the same item / chunk code as the per-pass kernels,
so results must equal the oracle's (after canonicalising the sign of zero)
and the per-call library path's byte for byte (signed zeros included), in
every queue configuration, on every graph replay, and with no dependency wait
abandoned."""

import ctypes

import numpy as np
import pytest

from adrt_amd import _lib, batched
from helpers import assert_bitwise
from oracle import oracle as O

gpu = pytest.mark.gpu


def _arr(specs):
    """adrt_hip_batch array of (batch, n, adrt?, bdrt?) with fake non-null
    pointers (the sizing and validation calls never dereference them)."""
    arr = (_lib.Batch * len(specs))()
    for e, (b, n, f, k) in zip(arr, specs):
        e.images, e.adrt, e.bdrt, e.batch, e.n = 256, (256 if f else None), (256 if k else None), b, n
    return arr


def test_sizes_without_device():
    lib = _lib.load()
    small = _arr([(7, 32, 1, 1), (3, 1, 1, 0)])
    assert lib.adrt_hip_queue_workspace_bytes(small, 2, _lib.F32) == 0  # LDS chunks only
    assert lib.adrt_hip_queue_state_bytes(small, 2) > 0
    one = _arr([(1, 256, 1, 1)])
    two = _arr([(2, 256, 1, 1)])
    w1 = lib.adrt_hip_queue_workspace_bytes(one, 1, _lib.F32)
    # one slot per image: the larger of the two intermediates, whole 256-byte lines
    assert w1 > 0 and w1 % 256 == 0
    assert lib.adrt_hip_queue_workspace_bytes(two, 1, _lib.F32) == 2 * w1
    assert lib.adrt_hip_queue_workspace_bytes(one, 1, _lib.F64) == 2 * w1
    # the slot holds at least the forward intermediate of one image
    assert w1 >= lib.adrt_hip_workspace_bytes(_lib.OP_ADRT, _lib.F32, 1, 256)
    # three completion counters per image of side >= 64
    s1, s2 = lib.adrt_hip_queue_state_bytes(one, 1), lib.adrt_hip_queue_state_bytes(two, 1)
    assert s2 - s1 == 3 * 4
    # invalid batches size to 0
    assert lib.adrt_hip_queue_workspace_bytes(_arr([(1, 2048, 1, 1)]), 1, _lib.F32) == 0
    assert lib.adrt_hip_queue_state_bytes(_arr([(1, 48, 1, 1)]), 1) == 0
    assert lib.adrt_hip_queue_state_bytes(one, 0) == 0


def test_plan_argument_checks():
    """adrt_hip_queue_plan validates before touching the device."""
    lib = _lib.load()
    plan = _lib.Queue()
    ok = _arr([(2, 128, 1, 1), (5, 16, 1, 0)])
    ws = lib.adrt_hip_queue_workspace_bytes(ok, 2, _lib.F32)
    sb = lib.adrt_hip_queue_state_bytes(ok, 2)

    def call(arr, count, dtype=_lib.F32, bundle=0, waves=0, ws_bytes=ws, state=256, sbytes=sb,
             p=ctypes.byref(plan)):
        return lib.adrt_hip_queue_plan(arr, count, dtype, bundle, waves, 256, ws_bytes, state,
                                       sbytes, p)

    assert call(ok, 2, p=None) == _lib.ERR_NULL
    assert call(ok, 2, dtype=5) == _lib.ERR_DTYPE
    assert call(ok, 0) == _lib.ERR_ARG
    assert call(_arr([(1, 8, 1, 1)] * 17), 17) == _lib.ERR_ARG
    assert call(ok, 2, bundle=3) == _lib.ERR_ARG
    assert call(ok, 2, waves=6) == _lib.ERR_ARG
    assert call(ok, 2, dtype=_lib.F64, waves=8) == _lib.ERR_ARG  # 8 waves: float32 only
    assert call(_arr([(1, 2048, 1, 1)]), 1) == _lib.ERR_SHAPE
    assert call(_arr([(1, 96, 1, 1)]), 1) == _lib.ERR_SHAPE
    assert call(_arr([(1, 128, 0, 1)]), 1) == _lib.ERR_NULL  # bdrt reads the forward output
    assert call(_arr([(1, 16, 0, 0)]), 1) == _lib.ERR_ARG
    assert call(_arr([(1 << 21, 16, 1, 1)]), 1) == _lib.ERR_TOO_BIG
    assert call(ok, 2, ws_bytes=ws - 1) == _lib.ERR_WORKSPACE
    assert call(ok, 2, sbytes=sb - 1) == _lib.ERR_WORKSPACE
    assert call(ok, 2, state=None) == _lib.ERR_NULL
    null = ctypes.cast(None, ctypes.POINTER(_lib.Batch))
    assert call(null, 1) == _lib.ERR_NULL
    assert lib.adrt_hip_queue_run(None, 0, None) == _lib.ERR_NULL
    bad = _lib.Queue()
    bad.state, bad.dtype, bad.entries, bad.lds_bytes, bad.waves = 256, _lib.F32, 1, 1024, 6
    assert lib.adrt_hip_queue_run(ctypes.byref(bad), 0, None) == _lib.ERR_ARG
    bad.waves = 8
    assert lib.adrt_hip_queue_run(ctypes.byref(bad), -1, None) == _lib.ERR_ARG
    bad.entries = 0
    assert lib.adrt_hip_queue_run(ctypes.byref(bad), 0, None) == _lib.OK  # nothing to launch


def _bytes_equal(a, b):
    import torch

    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    view = torch.int64 if a.dtype == torch.float64 else torch.int32
    return torch.equal(a.contiguous().view(view), b.contiguous().view(view))


@gpu
class TestQueue:
    SIDES = [1024, 512, 256, 128, 64, 32, 16, 8, 4, 2, 1]
    COUNTS = [2, 3, 2, 5, 3, 6, 4, 3, 2, 5, 3]

    def _stacks(self, cuda, dtype, sides=None, counts=None, seed=11):
        import torch

        g = torch.Generator(device=cuda)
        g.manual_seed(seed)
        return [torch.rand((b, n, n), generator=g, device=cuda, dtype=dtype)
                for n, b in zip(sides or self.SIDES, counts or self.COUNTS)]

    @pytest.mark.parametrize("dt", ["float32", "float64"])
    def test_eleven_sides_match_oracle(self, cuda, dt):
        import torch

        dtype = getattr(torch, dt)
        stacks = self._stacks(cuda, dtype)
        print(f"queue: sides {self.SIDES} counts {self.COUNTS} {dt}", flush=True)
        got = batched.adrt_bdrt_queue(stacks)
        for x, (f, k) in zip(stacks, got):
            xs = x.cpu().numpy()
            want_f = O.adrt(xs)
            assert_bitwise(f.cpu().numpy(), want_f, f"adrt n={x.shape[-1]}")
            assert_bitwise(k.cpu().numpy(), O.bdrt(want_f), f"bdrt n={x.shape[-1]}")

    def test_uneven_mix_matches_per_call_path(self, cuda):
        """1000 images with p(n) ~ 1/n over 32 .. 1024, every configuration,
        byte-equal (signed zeros included) to adrt / bdrt called per side."""
        import torch
        from adrt_amd import _kernels

        rng = np.random.default_rng(0)
        sides = np.array([32, 64, 128, 256, 512, 1024])
        p = (1.0 / sides) / (1.0 / sides).sum()
        draw = rng.choice(sides, size=1000, p=p)
        counts = [int((draw == n).sum()) for n in sides[::-1]]
        stacks = self._stacks(cuda, torch.float32, list(sides[::-1]), counts, seed=3)
        want = []
        for x in stacks:
            f = _kernels.adrt(x)
            want.append((f, _kernels.bdrt(f)))
        for bundle, waves in batched.UnevenBatch.QUEUE_CONFIGS:
            print(f"queue mix: counts {counts} bundle {bundle} waves {waves}", flush=True)
            got = batched.adrt_bdrt_queue(stacks, bundle=bundle, waves=waves)
            for (f, k), (wf, wk), n in zip(got, want, sides[::-1]):
                assert _bytes_equal(f, wf), (bundle, waves, n)
                assert _bytes_equal(k, wk), (bundle, waves, n)

    def test_forward_only_and_bundle_1(self, cuda):
        import torch
        from adrt_amd import _kernels

        stacks = self._stacks(cuda, torch.float32, [256, 64, 8], [3, 2, 4], seed=4)
        got = batched.adrt_bdrt_queue(stacks, bdrt=False, bundle=1)
        for x, (f, k) in zip(stacks, got):
            assert k is None and _bytes_equal(f, _kernels.adrt(x))

    def test_graph_replays(self, cuda):
        import torch
        from adrt_amd import _kernels

        sides = [512, 128, 64, 32, 2]
        counts = [2, 3, 1, 4, 3]
        plan = batched.UnevenBatch([n for n, b in zip(sides, counts) for _ in range(b)],
                                   autotune=False, mode="queue")
        assert plan.queue is not None
        g = torch.Generator(device=cuda)
        g.manual_seed(9)
        imgs = [torch.rand((n, n), generator=g, device=cuda) for n in plan.sizes]
        plan.load(imgs)
        plan.capture()
        eager_f = [_kernels.adrt(im.contiguous()) for im in imgs]
        eager_b = [_kernels.bdrt(f) for f in eager_f]
        for rep in range(4):
            for n in plan.classes:
                plan.fwd[n].fill_(np.nan)
                plan.back[n].fill_(np.nan)
            plan.replay()
            torch.cuda.synchronize()
            _ = _kernels.adrt(imgs[0].contiguous())  # eager work between replays
            for i in range(len(imgs)):
                assert _bytes_equal(plan.forward(i), eager_f[i]), (rep, i)
                assert _bytes_equal(plan.backprojection(i), eager_b[i]), (rep, i)
        assert plan.queue.timeouts() == 0

    def test_capped_grid(self, cuda):
        """A grid of 8 work-groups takes every entry in turn."""
        import torch
        from adrt_amd import _kernels

        stacks = self._stacks(cuda, torch.float32, [256, 128, 16], [2, 3, 5], seed=6)
        entries = []
        for x in stacks:
            b, n = x.shape[0], x.shape[-1]
            shape = (b, 4, 2 * n - 1, n)
            entries.append((x, torch.empty(shape, device=cuda), torch.empty(shape, device=cuda)))
        plan = batched.QueuePlan(entries)
        for _ in range(2):
            plan.run(max_groups=8)
        torch.cuda.synchronize()
        assert plan.timeouts() == 0
        for x, f, k in entries:
            wf = _kernels.adrt(x)
            assert _bytes_equal(f, wf) and _bytes_equal(k, _kernels.bdrt(wf))

    def test_auto_mode_records_timings(self, cuda):
        plan = batched.UnevenBatch([256, 256, 64, 32, 32, 8], reps=1)
        assert plan.queue_timings and plan.streams_ms is not None
        assert {(t["bundle"], t["waves"]) for t in plan.queue_timings} == \
            set(batched.UnevenBatch.QUEUE_CONFIGS)

    def test_rejects_bad_stacks(self, cuda):
        import torch

        x = torch.rand((1, 2048, 2048), device=cuda)
        with pytest.raises(ValueError, match="<= 1024"):
            batched.adrt_bdrt_queue([x])
        with pytest.raises(ValueError, match="at most 16"):
            batched.UnevenBatch([2048, 8], autotune=False, mode="queue")
        with pytest.raises(ValueError, match="expected 1 to 16"):
            batched.QueuePlan([])
