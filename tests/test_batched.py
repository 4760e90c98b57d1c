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


def test_candidates_outside_fused_range():
    # below 64 the default is the one-launch LDS kernel; above 4096 streaming only
    assert batched.candidates(_lib.OP_ADRT, 32, 100, 1) == [(_lib.VARIANT_STREAMING, 0),
                                                           (_lib.VARIANT_DEFAULT, 0)]
    assert batched.candidates(_lib.OP_BDRT, 8192, 2, 1) == [(_lib.VARIANT_STREAMING, 0)]
    at64 = batched.candidates(_lib.OP_ADRT, 64, 100, 1)
    assert (_lib.VARIANT_DEFAULT, 0) in at64 and (_lib.VARIANT_STREAMING, 0) in at64
    assert any(v == _lib.VARIANT_ALT_SPLIT for v, _ in at64)  # the fused kernels
    assert all(v != _lib.VARIANT_SINGLE for v, _ in batched.candidates(_lib.OP_ADRT, 128, 9, 1))


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
        assert lib.adrt_hip_set_config(_lib.OP_ADRT, 64, 4, 0) == _lib.ERR_ARG
        # the one-launch LDS kernel exists for n <= 64 and not for the inverse
        assert lib.adrt_hip_set_config(_lib.OP_ADRT, 128, _lib.VARIANT_SINGLE, 0) == _lib.ERR_ARG
        assert lib.adrt_hip_set_config(_lib.OP_IADRT, 32, _lib.VARIANT_SINGLE, 0) == _lib.ERR_ARG
        assert lib.adrt_hip_set_config(_lib.OP_BDRT, 64, _lib.VARIANT_SINGLE, 0) == _lib.OK
        assert lib.adrt_hip_workspace_bytes(_lib.OP_BDRT, _lib.F32, 4, 64) == 0

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

    @pytest.mark.parametrize("n,dt", [(n, "float32") for n in (1, 2, 16, 32, 64, 128, 512, 2048)]
                             + [(n, "float64") for n in (128, 512, 2048)])
    def test_every_variant_matches_oracle(self, cuda, reset_config, n, dt):
        """Every configuration the tuner can select (both fused splits, the
        streaming path, the LDS kernel, group budgets) gives the oracle's
        bytes for the forward and the back-projection; in float64 the
        default forward split differs from float32's and pass 1 at 2048
        with ceil(L/2) stages needs 160 KB of LDS."""
        import torch
        from adrt_amd import _kernels

        print(f"every variant n={n} {dt}", flush=True)
        b = 3 if n < 2048 else 1
        x = np.random.default_rng(n).standard_normal((b, n, n)).astype(dt)
        want = O.adrt(x)
        back_want = O.bdrt(want)
        xt = torch.from_numpy(x).to(cuda)
        code = _lib.F32 if dt == "float32" else _lib.F64
        per = max(int(_lib.load().adrt_hip_workspace_bytes(_lib.OP_ADRT, code, 1, n)), 1)
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
    plan.small, plan.large = [32, 64], [128, 256, 512, 1024]
    plan.code = _lib.F32
    plan.count = {32: 479, 64: 267, 128: 142, 256: 65, 512: 32, 1024: 15}
    ms = {128: (0.1, 0.15), 256: (0.15, 0.3), 512: (0.3, 0.5), 1024: (0.5, 1.0)}
    plan.tuning = [{"op": op, "n": n, "choice": {"ms": ms[n][i]}}
                   for n in plan.large for i, op in enumerate(("adrt", "bdrt"))]
    cost = dict((n, sum(ms[n])) for n in plan.large)
    plan.small_ms = 0.17                     # timed duration of the small launch
    cost["small"] = plan._class_cost("small")
    assert cost["small"] == 0.17
    lanes = plan._assign_lanes()
    assert len(lanes) == 4
    units = [u for lane in lanes for u in lane]
    assert sorted(map(str, units)) == sorted(map(str, plan.large + ["small"]))
    assert [1024] in lanes                      # the costliest class runs alone
    costs = [sum(cost[u] for u in lane) for lane in lanes]
    assert max(costs) == cost[1024]
    for lane in lanes:   # the small launch first, then the costliest first
        rest = [u for u in lane if u != "small"]
        assert lane[:len(lane) - len(rest)] in ([], ["small"])
        assert [cost[u] for u in rest] == sorted((cost[u] for u in rest), reverse=True)
    # untuned: bytes plus launch latency, same properties
    plan.tuning, plan.small_ms = [], None
    assert 0 < plan._class_cost("small") < plan._class_cost(1024)
    lanes = plan._assign_lanes()
    units = [u for lane in lanes for u in lane]
    assert sorted(map(str, units)) == sorted(map(str, plan.large + ["small"])) and len(lanes) == 4


def test_small_launch_argument_checks():
    """adrt_hip_adrt_bdrt_small validates before touching the device."""
    import ctypes

    lib = _lib.load()
    arr = (_lib.Batch * 9)()
    for e in arr:
        e.images, e.adrt, e.bdrt, e.batch, e.n = 16, 16, 16, 0, 32
    assert lib.adrt_hip_adrt_bdrt_small(arr, 0, _lib.F32, 0, None) == _lib.ERR_ARG
    assert lib.adrt_hip_adrt_bdrt_small(arr, 9, _lib.F32, 0, None) == _lib.ERR_ARG
    assert lib.adrt_hip_adrt_bdrt_small(arr, 1, 7, 0, None) == _lib.ERR_DTYPE
    assert lib.adrt_hip_adrt_bdrt_small(arr, 1, _lib.F32, -1, None) == _lib.ERR_ARG
    assert lib.adrt_hip_adrt_bdrt_small(ctypes.cast(None, ctypes.POINTER(_lib.Batch)), 1,
                                        _lib.F32, 0, None) == _lib.ERR_NULL
    arr[1].n = 128
    assert lib.adrt_hip_adrt_bdrt_small(arr, 2, _lib.F32, 0, None) == _lib.ERR_SHAPE
    arr[1].n = 48
    assert lib.adrt_hip_adrt_bdrt_small(arr, 2, _lib.F32, 0, None) == _lib.ERR_SHAPE
    arr[1].n, arr[1].adrt, arr[1].bdrt = 8, None, None
    assert lib.adrt_hip_adrt_bdrt_small(arr, 2, _lib.F32, 0, None) == _lib.ERR_ARG
    arr[1].adrt, arr[1].batch, arr[1].images = 16, 3, None
    assert lib.adrt_hip_adrt_bdrt_small(arr, 2, _lib.F32, 0, None) == _lib.ERR_NULL
    arr[1].images, arr[1].batch = 16, 1 << 40
    assert lib.adrt_hip_adrt_bdrt_small(arr, 2, _lib.F32, 0, None) == _lib.ERR_TOO_BIG
    # nothing to do: every batch empty
    assert lib.adrt_hip_adrt_bdrt_small(arr, 1, _lib.F32, 0, None) == _lib.OK


def _same_signed(a, b):
    """Bitwise equal including signs of zero; NaNs match by position (their
    payload depends on the order of two NaN operands)."""
    na, nb = np.isnan(a), np.isnan(b)
    return (a.shape == b.shape and np.array_equal(na, nb)
            and np.array_equal(np.where(na, 0, a).view(np.uint8), np.where(nb, 0, b).view(np.uint8)))


@gpu
class TestSmallLaunch:
    """All sides <= 64 in one persistent launch of the LDS kernel."""

    SIDES = [64, 32, 16, 8, 4, 2, 1]
    COUNTS = [5, 9, 3, 7, 2, 3, 4]

    def _stacks(self, cuda, dtype):
        import torch

        g = torch.Generator(device=cuda)
        g.manual_seed(5)
        return [torch.randn((b, n, n), generator=g, device=cuda, dtype=dtype)
                for n, b in zip(self.SIDES, self.COUNTS)]

    @pytest.mark.parametrize("dt", ["float32", "float64"])
    def test_matches_oracle_and_default_path(self, cuda, dt):
        import torch
        from adrt_amd import _kernels

        stacks = self._stacks(cuda, getattr(torch, dt))
        stacks[0][1, 2, :3] = torch.tensor([-0.0, float("inf"), -float("inf")], dtype=stacks[0].dtype)
        got = batched.adrt_bdrt_small(stacks)
        for x, (f, k) in zip(stacks, got):
            xn = x.cpu().numpy()
            want = O.adrt(xn)
            assert_bitwise(f.cpu().numpy(), want, f"adrt {tuple(x.shape)}")
            assert_bitwise(k.cpu().numpy(), O.bdrt(want), f"bdrt {tuple(x.shape)}")
            # the LDS kernel also keeps the oracle's signs of zero (NaNs by position)
            assert _same_signed(f.cpu().numpy(), want), f"adrt {tuple(x.shape)}"
            assert _same_signed(k.cpu().numpy(), O.bdrt(want)), f"bdrt {tuple(x.shape)}"
            # the library's single-op calls (fused kernels at 64) agree bit for
            # bit (inf - inf makes NaNs here, which torch.equal calls unequal)
            assert_bitwise(f.cpu().numpy(), _kernels.adrt(x).cpu().numpy())
            assert_bitwise(k.cpu().numpy(), _kernels.bdrt(f).cpu().numpy())

    def test_one_output(self, cuda):
        import torch

        stacks = self._stacks(cuda, torch.float32)[:3]
        both = batched.adrt_bdrt_small(stacks)
        only_f = batched.adrt_bdrt_small(stacks, bdrt=False)
        only_b = batched.adrt_bdrt_small(stacks, adrt=False)
        for (f, k), (f1, k1), (f2, k2) in zip(both, only_f, only_b):
            assert k1 is None and f2 is None
            assert torch.equal(f, f1) and torch.equal(k, k2)

    def test_graph_replays(self, cuda):
        import torch

        stacks = self._stacks(cuda, torch.float32)
        entries = []
        for x in stacks:
            b, n = x.shape[0], x.shape[-1]
            entries.append((x, torch.empty((b, 4, 2 * n - 1, n), device=cuda),
                            torch.empty((b, 4, 2 * n - 1, n), device=cuda)))
        arr = batched._batch_array(entries)
        want = batched.adrt_bdrt_small(stacks)
        s = torch.cuda.Stream(cuda)
        s.wait_stream(torch.cuda.current_stream(cuda))
        with torch.cuda.stream(s):
            batched._launch_small(arr, _lib.F32, s)
        torch.cuda.current_stream(cuda).wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            batched._launch_small(arr, _lib.F32, torch.cuda.current_stream(cuda))
        for rep in range(3):
            for _, f, k in entries:
                f.fill_(float("nan"))
                k.fill_(float("nan"))
            g.replay()
            torch.cuda.synchronize()
            _ = batched.adrt_bdrt_small(stacks[:1])  # eager work between replays
            for (_, f, k), (wf, wk) in zip(entries, want):
                assert torch.equal(f, wf) and torch.equal(k, wk), rep

    def test_capped_grid_strides_over_chunks(self, cuda):
        """A grid capped at 32 work-groups runs ~15 chunks per work-group
        through the persistent loop (and the slot permutation's partial last
        group): the same bytes as the default grid and as the oracle."""
        import torch

        g = torch.Generator(device=cuda)
        g.manual_seed(9)
        stacks = [torch.randn((101, 64, 64), generator=g, device=cuda),
                  torch.randn((37, 32, 32), generator=g, device=cuda),
                  torch.randn((3, 8, 8), generator=g, device=cuda)]
        want = batched.adrt_bdrt_small(stacks)
        entries = []
        for x in stacks:
            b, n = x.shape[0], x.shape[-1]
            entries.append((x, torch.full((b, 4, 2 * n - 1, n), float("nan"), device=cuda),
                            torch.full((b, 4, 2 * n - 1, n), float("nan"), device=cuda)))
        arr = batched._batch_array(entries)
        batched._launch_small(arr, _lib.F32, torch.cuda.current_stream(cuda), max_groups=32)
        torch.cuda.synchronize()
        for (x, f, k), (wf, wk) in zip(entries, want):
            assert torch.equal(f, wf) and torch.equal(k, wk), tuple(x.shape)
        xn = stacks[1][:2].cpu().numpy()
        assert _same_signed(entries[1][1][:2].cpu().numpy(), O.adrt(xn))

    def test_variant_single_at_64(self, cuda, reset_config):
        """adrt/bdrt at n = 64 configured to the LDS kernel: no workspace and
        the fused path's results."""
        import torch
        from adrt_amd import _kernels

        x = self._stacks(cuda, torch.float32)[0]
        f0 = _kernels.adrt(x)
        k0 = _kernels.bdrt(f0)
        batched.set_config(_lib.OP_ADRT, 64, _lib.VARIANT_SINGLE, 0)
        batched.set_config(_lib.OP_BDRT, 64, _lib.VARIANT_SINGLE, 0)
        assert _lib.load().adrt_hip_workspace_bytes(_lib.OP_ADRT, _lib.F32, 5, 64) == 0
        assert torch.equal(_kernels.adrt(x), f0)
        assert torch.equal(_kernels.bdrt(f0), k0)

    def test_rejects_bad_stacks(self, cuda):
        import torch

        with pytest.raises(ValueError, match="power of two"):
            batched.adrt_bdrt_small([torch.zeros((1, 128, 128), device=cuda)])
        with pytest.raises(ValueError, match="one dtype"):
            batched.adrt_bdrt_small([torch.zeros((1, 8, 8), device=cuda),
                                     torch.zeros((1, 4, 4), device=cuda, dtype=torch.float64)])
        with pytest.raises(ValueError, match="1 to 8"):
            batched.adrt_bdrt_small([torch.zeros((1, 8, 8), device=cuda)] * 9)


def test_tune_cache_file(tmp_path):
    """The cache round-trips entries, survives a corrupt file and writes
    atomically (no temporary files left behind)."""
    path = tmp_path / "sub" / "tune.json"
    c = batched.TuneCache(str(path))
    assert c.get("k") is None  # no file yet
    c.put("k", {"variant": 1, "group_bytes": 64 * MIB, "ms": 0.5})
    c.put("j", {"pipeline": "streams"})
    assert batched.TuneCache(str(path)).get("k") == {"variant": 1, "group_bytes": 64 * MIB,
                                                      "ms": 0.5}
    assert batched.TuneCache(str(path)).get("j") == {"pipeline": "streams"}
    assert sorted(p.name for p in path.parent.iterdir()) == ["tune.json"]
    path.write_text("{not json")
    assert c.get("k") is None
    c.put("k", {"variant": 0, "group_bytes": 0})
    assert c.get("k") == {"variant": 0, "group_bytes": 0}


def test_tune_cache_path_from_environment(tmp_path, monkeypatch):
    monkeypatch.setenv("ADRT_AMD_TUNE_CACHE", str(tmp_path / "t.json"))
    assert batched.TuneCache().path == str(tmp_path / "t.json")
    monkeypatch.delenv("ADRT_AMD_TUNE_CACHE")
    assert batched.TuneCache().path.endswith("tune.json")


@gpu
class TestTuneCache:
    def test_tune_reuses_cached_choice(self, cuda, reset_config, tmp_path):
        c = batched.TuneCache(str(tmp_path / "tune.json"))
        first = batched.tune(_lib.OP_BDRT, 256, 6, reps=1, cache=c)
        assert not first["cached"] and first["timings"]
        batched.set_config(_lib.OP_BDRT, 256, _lib.VARIANT_STREAMING, 0)
        again = batched.tune(_lib.OP_BDRT, 256, 6, reps=1, cache=c)
        assert again["cached"] and again["timings"] == []
        want = (first["choice"]["variant"],
                first["choice"]["group_bytes"] or 136 * MIB)
        assert batched.get_config(_lib.OP_BDRT, 256) == want
        other = batched.tune(_lib.OP_BDRT, 256, 7, reps=1, cache=c)  # another batch size
        assert not other["cached"]

    def test_uneven_plan_reuses_pipeline_choice(self, cuda, reset_config, tmp_path):
        c = batched.TuneCache(str(tmp_path / "tune.json"))
        sizes = [128, 128, 64, 32, 32, 8]
        first = batched.UnevenBatch(sizes, reps=1, cache=c)
        assert not first.queue_cached and first.queue_timings
        second = batched.UnevenBatch(sizes, reps=1, cache=c)
        assert second.queue_cached and not second.queue_timings
        assert (second.queue is None) == (first.queue is None)
        assert all(t["cached"] for t in second.tuning)
