"""Forward ADRT and back-projection of many images of uneven sizes.

``UnevenBatch`` takes a list of image sides (powers of two) and holds, per
distinct side n, one stacked device buffer of the images of that side, their
forward ADRTs and back-projections, and the workspaces.  ``run()`` enqueues
all sides n <= 64 as one persistent launch of the LDS kernel (every stage of
both transforms in LDS, ``adrt_hip_adrt_bdrt_small``) and, per larger side,
one batched forward call and one batched back-projection call; these units
go onto at most 4 streams.  Alternatively (``mode="queue"``, or chosen by
the tuner under ``mode="auto"``) the whole batch runs as ONE persistent
launch with a stage-ordered work queue (``QueuePlan``, sides up to 1024):
forward pass 1 of every image, the small sides' LDS chunks, forward pass 2,
then back-projection passes A and B, each entry waiting only for its own
image's previous pass.  ``capture()`` records either pipeline once as a
hipGraph that ``replay()`` relaunches with no host work.

``tune()`` is the autotuner: for one (op, n, batch) it times every kernel
variant the library offers for that size (the one-launch LDS kernel for
n <= 64, the fused two-pass kernels with either stage split, i.e. the number
of stages fused in LDS per pass, and the streaming one-kernel-per-stage
path) and every launch-group size that makes a difference, with HIP events on data of that shape, then leaves the fastest
configuration selected in the library.  Results are bit-identical for every
configuration; only speed differs.

The configuration is process-wide per (op, n) (``adrt_hip_set_config``): a
later ``tune`` of the same side for another batch size changes what eager
calls of an earlier plan use (a captured graph keeps the kernels it
recorded).  Plans check their workspaces against the library on every eager
call and raise if the configuration changed under them.
"""

import ctypes
import json
import os
import tempfile

from . import _device as dev
from . import _lib

_MIB = 1 << 20
# Launch-group budgets tried by the tuner (bytes of intermediate per group).
GROUP_BUDGETS = (16 * _MIB, 64 * _MIB, 136 * _MIB, 512 * _MIB)


def _torch():
    return dev.torch()


def _code(dtype):
    torch = _torch()
    if dtype == torch.float32:
        return _lib.F32
    if dtype == torch.float64:
        return _lib.F64
    raise TypeError(f"unsupported array dtype {dtype}")


def get_config(op, n):
    """(variant, group_bytes) in effect for (op, n)."""
    lib = _lib.load()
    v, g = ctypes.c_int(), ctypes.c_size_t()
    _lib.check(lib.adrt_hip_get_config(op, n, ctypes.byref(v), ctypes.byref(g)), "get_config")
    return int(v.value), int(g.value)


def set_config(op, n, variant, group_bytes=0):
    _lib.check(_lib.load().adrt_hip_set_config(op, n, int(variant), int(group_bytes)),
               "set_config")


def fused_range(n):
    L = int(n).bit_length() - 1
    return 6 <= L <= 12


SMALL_MAX = 64  # sides served by the one-launch LDS kernel
SMALL_CLASSES = 8  # size classes per launch of that kernel


def small_range(n):
    return 1 <= n <= SMALL_MAX


def _batch_array(entries):
    """ctypes array of adrt_hip_batch for (images, adrt, bdrt) tensor triples."""
    arr = (_lib.Batch * len(entries))()
    for e, (img, fwd, back) in zip(arr, entries):
        e.images = img.data_ptr()
        e.adrt = fwd.data_ptr() if fwd is not None else None
        e.bdrt = back.data_ptr() if back is not None else None
        e.batch = img.shape[0]
        e.n = img.shape[-1]
    return arr


def _launch_small(arr, code, stream, max_groups=0):
    _lib.check(_lib.load().adrt_hip_adrt_bdrt_small(arr, len(arr), code, int(max_groups),
                                                    ctypes.c_void_p(stream.cuda_stream)),
               "adrt_bdrt_small")


def adrt_bdrt_small(stacks, *, adrt=True, bdrt=True):
    """Forward ADRT and back-projection of several stacks of small images in
    one persistent launch.

    ``stacks``: up to 8 tensors (B, n, n), n a power of two <= 64, one dtype
    (float32 or float64) and device.  Returns one (adrt, bdrt) pair per stack,
    each (B, 4, 2n-1, n) or None where not asked; equal bit for bit to
    ``adrt(x)`` and ``bdrt(adrt(x))``."""
    torch = _torch()
    dev.require_gpu()
    if not stacks or len(stacks) > SMALL_CLASSES:
        raise ValueError(f"expected 1 to {SMALL_CLASSES} stacks, got {len(stacks)}")
    if not (adrt or bdrt):
        raise ValueError("ask for the forward result, the back-projection or both")
    dtype, device = stacks[0].dtype, stacks[0].device
    code = _code(dtype)
    entries = []
    for x in stacks:
        if x.dim() != 3 or x.shape[1] != x.shape[2] or not small_range(x.shape[-1]) \
                or x.shape[-1] & (x.shape[-1] - 1):
            raise ValueError(f"stack of shape {tuple(x.shape)}: expected (B, n, n), "
                             f"n a power of two <= {SMALL_MAX}")
        if x.dtype != dtype or x.device != device or x.device.type != "cuda":
            raise ValueError("stacks must share one dtype and one GPU device")
        x = x.contiguous()
        b, n = x.shape[0], x.shape[-1]
        shape = (b, 4, 2 * n - 1, n)
        f = torch.empty(shape, dtype=dtype, device=device) if adrt else None
        k = torch.empty(shape, dtype=dtype, device=device) if bdrt else None
        entries.append((x, f, k))
    with torch.cuda.device(device):
        _launch_small(_batch_array(entries), code, torch.cuda.current_stream(device))
    return [(f, k) for _, f, k in entries]


QUEUE_MAX = 1024  # largest side of the persistent queue kernel
QUEUE_BATCHES = 16  # entries per queue


class QueuePlan:
    """Forward ADRT and back-projection of several stacks of images of any
    power-of-two sides n <= 1024 in ONE persistent launch
    (``adrt_hip_queue_run``): a stage-ordered work queue of the fused passes
    of every image of side >= 64 and the LDS chunks of the sides <= 32.

    ``entries``: (images (B, n, n), adrt (B, 4, 2n-1, n), bdrt or None)
    tensor triples, one dtype and device; the tensors must stay alive and in
    place while the plan is used.  ``bundle``: fused items per queue entry
    (1, 2, 4, 8); ``waves``: register budget of the kernel (8 or 4 waves per
    SIMD; 0 = the dtype's default).  The plan owns its device workspace and
    state; ``run()`` is one launch, capturable in a graph and replayable."""

    def __init__(self, entries, *, bundle=8, waves=0):
        torch = _torch()
        dev.require_gpu()
        if not entries or len(entries) > QUEUE_BATCHES:
            raise ValueError(f"expected 1 to {QUEUE_BATCHES} stacks, got {len(entries)}")
        self.dtype, self.device = entries[0][0].dtype, entries[0][0].device
        self.code = _code(self.dtype)
        for img, fwd, back in entries:
            n = img.shape[-1]
            if img.dim() != 3 or img.shape[1] != n or n > QUEUE_MAX or n & (n - 1):
                raise ValueError(f"stack of shape {tuple(img.shape)}: expected (B, n, n), "
                                 f"n a power of two <= {QUEUE_MAX}")
            for x in (img, fwd, back):
                if x is None:
                    continue
                if x.dtype != self.dtype or x.device != self.device or not x.is_contiguous():
                    raise ValueError("stacks must be contiguous and share one dtype and device")
            out = (img.shape[0], 4, 2 * n - 1, n)
            if fwd is None or tuple(fwd.shape) != out or (back is not None
                                                          and tuple(back.shape) != out):
                raise ValueError(f"outputs of shape {out} expected for stack {tuple(img.shape)}")
        self.entries = list(entries)
        self.batches = _batch_array(self.entries)
        lib = _lib.load()
        nb = len(self.entries)
        ws = int(lib.adrt_hip_queue_workspace_bytes(self.batches, nb, self.code))
        sb = int(lib.adrt_hip_queue_state_bytes(self.batches, nb))
        if sb == 0:
            raise ValueError("invalid batches for the queue")
        self.workspace = torch.empty(max(ws, 1), dtype=torch.uint8, device=self.device)
        self.state = torch.empty(sb, dtype=torch.uint8, device=self.device)
        self.plan = _lib.Queue()
        with torch.cuda.device(self.device):
            _lib.check(lib.adrt_hip_queue_plan(self.batches, nb, self.code, int(bundle),
                                               int(waves), self.workspace.data_ptr(), ws,
                                               self.state.data_ptr(), sb,
                                               ctypes.byref(self.plan)), "queue_plan")
        self.bundle, self.waves = int(bundle), int(self.plan.waves)

    def run(self, stream=None, max_groups=0):
        torch = _torch()
        stream = stream or torch.cuda.current_stream(self.device)
        _lib.check(_lib.load().adrt_hip_queue_run(ctypes.byref(self.plan), int(max_groups),
                                                  ctypes.c_void_p(stream.cuda_stream)),
                   "queue_run")

    def timeouts(self):
        """Dependency waits abandoned since the last call (0 unless the
        device stalled for seconds; nonzero means invalid results)."""
        c = ctypes.c_int()
        with _torch().cuda.device(self.device):
            _lib.check(_lib.load().adrt_hip_queue_timeouts(ctypes.byref(self.plan),
                                                           ctypes.byref(c)), "queue_timeouts")
        return int(c.value)


def adrt_bdrt_queue(stacks, *, bdrt=True, bundle=8, waves=0):
    """Forward ADRT (and back-projection) of up to 16 stacks (B, n, n) of
    power-of-two sides n <= 1024 in one persistent launch.  Returns one
    (adrt, bdrt) pair per stack (bdrt None when not asked), equal bit for bit
    to ``adrt(x)`` and ``bdrt(adrt(x))``."""
    torch = _torch()
    dev.require_gpu()
    entries = []
    for x in stacks:
        if x.dim() != 3:
            raise ValueError(f"stack of shape {tuple(x.shape)}: expected (B, n, n)")
        x = x.contiguous()
        b, n = x.shape[0], x.shape[-1]
        shape = (b, 4, 2 * n - 1, n)
        f = torch.empty(shape, dtype=x.dtype, device=x.device)
        k = torch.empty(shape, dtype=x.dtype, device=x.device) if bdrt else None
        entries.append((x, f, k))
    plan = QueuePlan(entries, bundle=bundle, waves=waves)
    with torch.cuda.device(plan.device):
        plan.run()
    torch.cuda.current_stream(plan.device).synchronize()
    if plan.timeouts():
        raise RuntimeError("the queue kernel abandoned a dependency wait; results invalid")
    return [(f, k) for _, f, k in entries]


def candidates(op, n, batch, per_image_bytes):
    """Distinct (variant, group_bytes) configurations for (op, n, batch).

    The one-launch LDS kernel is the default for n <= 64 (the fused kernels
    are ALT_SPLIT there); the alternative stage split exists for odd log2 n
    in the fused range;
    group budgets are kept only when they change the images per launch group
    (``per_image_bytes``: the fused path's intermediate per image)."""
    L = int(n).bit_length() - 1
    out = [(_lib.VARIANT_STREAMING, 0)]
    if small_range(n):  # up to 64 the default is the LDS kernel
        out.append((_lib.VARIANT_DEFAULT, 0))
    if not fused_range(n):
        return out
    seen = set()
    budgets = []
    for b in GROUP_BUDGETS:
        g = min(max(b // per_image_bytes, 1), batch)
        if g not in seen:
            seen.add(g)
            budgets.append(b)
    if small_range(n):  # the fused kernels at 64: ALT_SPLIT (the same split)
        variants = [_lib.VARIANT_ALT_SPLIT]
    else:
        variants = [_lib.VARIANT_DEFAULT] + ([_lib.VARIANT_ALT_SPLIT] if L % 2 == 1 else [])
    out += [(v, b) for v in variants for b in budgets]
    return out


def _launch(op, src, dst, ws, code, batch, n, stream):
    lib = _lib.load()
    fn = lib.adrt_hip_adrt if op == _lib.OP_ADRT else lib.adrt_hip_bdrt
    _lib.check(fn(dev.ptr(src), dev.ptr(dst), dev.ptr(ws), ws.numel(), code, batch, n,
                  ctypes.c_void_p(stream.cuda_stream)),
               "adrt" if op == _lib.OP_ADRT else "bdrt")


def _workspace(op, code, batch, n, device):
    nbytes = _lib.load().adrt_hip_workspace_bytes(op, code, batch, n)
    return _torch().empty(max(int(nbytes), 1), dtype=_torch().uint8, device=device)


class TuneCache:
    """The autotuner's cache file: a JSON object of tuning results keyed by
    device name, library ABI version, dtype and workload, so that a later
    process selects the same configuration without timing again.

    Path: ``path``, else $ADRT_AMD_TUNE_CACHE, else
    ~/.cache/adrt_amd/tune.json.  An unreadable or corrupt file counts as
    empty; writes go through a temporary file and an atomic rename."""

    def __init__(self, path=None):
        self.path = path or os.environ.get("ADRT_AMD_TUNE_CACHE") or os.path.join(
            os.path.expanduser("~"), ".cache", "adrt_amd", "tune.json")

    @staticmethod
    def key(device, dtype, *parts):
        torch = _torch()
        name = torch.cuda.get_device_name(device)
        abi = int(_lib.load().adrt_hip_abi_version())
        return "|".join([name, f"abi{abi}", str(dtype).replace("torch.", "")]
                        + [str(p) for p in parts])

    def _read(self):
        try:
            with open(self.path) as f:
                data = json.load(f)
            return data if isinstance(data, dict) else {}
        except (OSError, ValueError):
            return {}

    def get(self, key):
        return self._read().get(key)

    def put(self, key, value):
        data = self._read()
        data[key] = value
        folder = os.path.dirname(os.path.abspath(self.path))
        os.makedirs(folder, exist_ok=True)
        fd, tmp = tempfile.mkstemp(dir=folder, prefix=".tune-", suffix=".json")
        try:
            with os.fdopen(fd, "w") as f:
                json.dump(data, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            if os.path.exists(tmp):
                os.unlink(tmp)


def tune(op, n, batch, *, dtype=None, device=None, reps=3, cache=None):
    """Times every candidate configuration of (op, n, batch) and selects the
    fastest in the library.  Returns a dict with the choice and all timings
    (milliseconds per call, mean of ``reps`` after one warm-up).

    ``cache``: a TuneCache (or None for none).  A cached choice for the
    same device, library, dtype and (op, n, batch) is selected without
    timing (the result then has ``"cached": True`` and no timings)."""
    torch = _torch()
    dev.require_gpu()
    dtype = dtype or torch.float32
    device = device or torch.device("cuda", torch.cuda.current_device())
    code = _code(dtype)
    name = "adrt" if op == _lib.OP_ADRT else "bdrt"
    key = TuneCache.key(device, dtype, name, n, batch) if cache is not None else None
    if key is not None:
        hit = cache.get(key)
        if isinstance(hit, dict) and "variant" in hit and "group_bytes" in hit:
            try:
                set_config(op, n, int(hit["variant"]), int(hit["group_bytes"]))
                return {"op": name, "n": n, "batch": batch, "choice": hit, "timings": [],
                        "cached": True}
            except (RuntimeError, ValueError):
                pass  # a configuration this library rejects: tune again
    set_config(op, n, _lib.VARIANT_DEFAULT, 0)
    per = max(int(_lib.load().adrt_hip_workspace_bytes(op, code, 1, n)), 1)
    if op == _lib.OP_ADRT:
        src = torch.rand((batch, n, n), dtype=dtype, device=device)
    else:
        src = torch.rand((batch, 4, 2 * n - 1, n), dtype=dtype, device=device)
    dst = torch.empty((batch, 4, 2 * n - 1, n), dtype=dtype, device=device)
    stream = torch.cuda.current_stream(device)
    timings = []
    for variant, budget in candidates(op, n, batch, per):
        set_config(op, n, variant, budget)
        ws = _workspace(op, code, batch, n, device)
        _launch(op, src, dst, ws, code, batch, n, stream)
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(reps):
            _launch(op, src, dst, ws, code, batch, n, stream)
        end.record(stream)
        end.synchronize()
        timings.append({"variant": variant, "group_bytes": budget,
                        "ms": start.elapsed_time(end) / reps})
        del ws
    best = min(timings, key=lambda t: t["ms"])
    set_config(op, n, best["variant"], best["group_bytes"])
    if key is not None:
        cache.put(key, best)
    return {"op": name, "n": n, "batch": batch, "choice": best, "timings": timings,
            "cached": False}


def _time_ms(fn, stream, reps):
    torch = _torch()
    fn()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        fn()
    end.record(stream)
    end.synchronize()
    return start.elapsed_time(end) / reps


class UnevenBatch:
    """Forward ADRT + back-projection plan for images of sides ``sizes``.

    ``mode``: "streams" (one forward and one back-projection call per side,
    on up to 4 streams), "queue" (one persistent launch of the stage-ordered
    work queue, ``QueuePlan``; sides <= 1024) or "auto" (with ``autotune``,
    whichever of the streams and the queue configurations measures fastest;
    without, the streams).  ``cache``: a TuneCache that keeps the per-side
    choices and the pipeline choice across processes (None: no cache)."""

    QUEUE_CONFIGS = ((8, 8), (4, 8), (8, 4))  # (bundle, waves) tried for float32

    def __init__(self, sizes, *, dtype=None, device=None, autotune=True, reps=3, mode="auto",
                 cache=None):
        torch = _torch()
        dev.require_gpu()
        self.dtype = dtype or torch.float32
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.code = _code(self.dtype)
        self.sizes = [int(s) for s in sizes]
        for s in self.sizes:
            if s < 1 or s & (s - 1):
                raise ValueError("array must be square with a power of two shape")
        self.classes = sorted(set(self.sizes))
        self.slot = []  # original index -> (n, position in its class)
        count = {n: 0 for n in self.classes}
        for s in self.sizes:
            self.slot.append((s, count[s]))
            count[s] += 1
        self.count = count
        # sides <= 64 run together in one launch (at most 7 classes: 1 .. 64)
        self.small = [n for n in self.classes if small_range(n)]
        self.large = [n for n in self.classes if not small_range(n)]
        self.tuning = []
        if autotune:
            for n in self.large:
                for op in (_lib.OP_ADRT, _lib.OP_BDRT):
                    self.tuning.append(tune(op, n, count[n], dtype=self.dtype,
                                            device=self.device, reps=reps, cache=cache))
        self.images, self.fwd, self.back, self.ws_f, self.ws_b = {}, {}, {}, {}, {}
        for n in self.classes:
            b = count[n]
            self.images[n] = torch.zeros((b, n, n), dtype=self.dtype, device=self.device)
            self.fwd[n] = torch.empty((b, 4, 2 * n - 1, n), dtype=self.dtype, device=self.device)
            self.back[n] = torch.empty_like(self.fwd[n])
            if n in self.large:
                self.ws_f[n] = _workspace(_lib.OP_ADRT, self.code, b, n, self.device)
                self.ws_b[n] = _workspace(_lib.OP_BDRT, self.code, b, n, self.device)
        self.small_batches = _batch_array(
            [(self.images[n], self.fwd[n], self.back[n]) for n in self.small]
        ) if self.small else None
        self.small_ms = None
        if autotune and self.small:  # time the small launch like the tuner
            stream = torch.cuda.current_stream(self.device)
            _launch_small(self.small_batches, self.code, stream)
            start, end = (torch.cuda.Event(enable_timing=True),
                          torch.cuda.Event(enable_timing=True))
            start.record(stream)
            for _ in range(reps):
                _launch_small(self.small_batches, self.code, stream)
            end.record(stream)
            end.synchronize()
            self.small_ms = start.elapsed_time(end) / reps
        self.lanes = self._assign_lanes()
        self.streams = [torch.cuda.Stream(self.device) for _ in self.lanes]
        self.graph = None
        if mode not in ("auto", "streams", "queue"):
            raise ValueError(f"unknown mode {mode!r}")
        self.queue = None
        self.queue_timings = []
        self.queue_cached = False
        self.streams_ms = None
        queue_ok = max(self.classes) <= QUEUE_MAX and len(self.classes) <= QUEUE_BATCHES
        if mode == "queue" and not queue_ok:
            raise ValueError(f"the queue takes at most {QUEUE_BATCHES} sides, each <= {QUEUE_MAX}")
        if mode == "queue" or (mode == "auto" and autotune and queue_ok):
            self._tune_queue(reps if autotune else 0, keep_streams=mode == "auto", cache=cache)

    def _queue_entries(self):
        return [(self.images[n], self.fwd[n], self.back[n]) for n in reversed(self.classes)]

    def _tune_queue(self, reps, keep_streams, cache=None):
        """Times the queue configurations (and the streams) and keeps the
        fastest; with reps == 0 takes the default queue configuration."""
        torch = _torch()
        stream = torch.cuda.current_stream(self.device)
        configs = self.QUEUE_CONFIGS if self.code == _lib.F32 else ((8, 4), (4, 4))
        if reps == 0:
            self.queue = QueuePlan(self._queue_entries())
            return
        key = None
        if cache is not None:
            mix = ",".join(f"{n}x{self.count[n]}" for n in self.classes)
            key = TuneCache.key(self.device, self.dtype, "uneven", mix,
                                "auto" if keep_streams else "queue")
            hit = cache.get(key)
            if isinstance(hit, dict) and hit.get("pipeline") in ("streams", "queue"):
                if hit["pipeline"] == "queue":
                    self.queue = QueuePlan(self._queue_entries(), bundle=int(hit["bundle"]),
                                           waves=int(hit["waves"]))
                self.queue_cached = True
                return
        best, best_ms = None, None
        for bundle, waves in configs:
            q = QueuePlan(self._queue_entries(), bundle=bundle, waves=waves)
            ms = _time_ms(lambda: q.run(stream), stream, reps)
            if q.timeouts():
                raise RuntimeError("the queue kernel abandoned a dependency wait")
            self.queue_timings.append({"bundle": bundle, "waves": waves, "ms": ms})
            if best_ms is None or ms < best_ms:
                best, best_ms = q, ms
            else:
                del q
        if keep_streams:
            self.streams_ms = _time_ms(self._enqueue_streams, stream, reps)
        if keep_streams and self.streams_ms <= best_ms:
            choice = {"pipeline": "streams", "ms": self.streams_ms}
        else:
            self.queue = best
            choice = {"pipeline": "queue", "bundle": best.bundle, "waves": best.waves,
                      "ms": best_ms}
        if key is not None:
            cache.put(key, choice)

    # The runtime maps streams onto 4 hardware queues (HIP's default), so
    # more than 4 streams share queues and serialise whatever lands
    # together.  The units (each large side, and the one launch of all small
    # sides) are spread over at most 4 streams by
    # longest-processing-time-first: each unit (costliest first) goes to the
    # stream with the least work so far.  A large side's cost is its tuned
    # forward + back-projection time, or, untuned, its bytes plus a fixed
    # per-launch latency; the small launch's cost is its timed duration
    # (untuned: its bytes at 1 TB/s plus one launch).
    HW_QUEUES = 4
    SMALL = "small"  # the unit of all sides <= 64
    SMALL_GROUPS = 0  # grid cap of the small launch (0: every resident work-group)

    def _bytes(self, n):
        elem = 4 if self.code == _lib.F32 else 8
        return self.count[n] * elem * (n * n + 3 * 4 * n * (2 * n - 1))

    def _class_cost(self, n):
        if n == self.SMALL:
            if getattr(self, "small_ms", None):
                return self.small_ms
            return sum(self._bytes(m) for m in self.small) / 1e9 + 0.01  # ms
        ms = [t["choice"]["ms"] for t in self.tuning if t["n"] == n and "ms" in t["choice"]]
        if len(ms) == 2:
            return sum(ms)
        launches = 4 if fused_range(n) else 2 * (n.bit_length() + 2)
        return self._bytes(n) / 3e9 + launches * 0.01  # ms

    def _assign_lanes(self):
        units = list(self.large) + ([self.SMALL] if self.small else [])
        lanes = [[] for _ in range(min(self.HW_QUEUES, len(units)))]
        load = [0.0] * len(lanes)
        for u in sorted(units, key=self._class_cost, reverse=True):
            i = min(range(len(lanes)), key=lambda j: load[j])
            lanes[i].append(u)
            load[i] += self._class_cost(u)
        # The small launch is bound by instruction issue, the large sides by
        # HBM bandwidth: it goes first on its stream, beside the large sides'
        # first kernels, rather than last and alone.
        return [[u for u in lane if u == self.SMALL] + [u for u in lane if u != self.SMALL]
                for lane in lanes]

    def load(self, images):
        """Copies ``images`` (a sequence of (n, n) tensors in plan order)."""
        if len(images) != len(self.sizes):
            raise ValueError(f"expected {len(self.sizes)} images, got {len(images)}")
        for img, (n, k) in zip(images, self.slot):
            if tuple(img.shape) != (n, n):
                raise ValueError(f"image of shape {tuple(img.shape)} where ({n}, {n}) was planned")
            self.images[n][k].copy_(img)

    def _enqueue(self):
        if self.queue is not None:
            self.queue.run(_torch().cuda.current_stream(self.device))
        else:
            self._enqueue_streams()

    def _enqueue_streams(self):
        torch = _torch()
        main = torch.cuda.current_stream(self.device)
        if len(self.lanes) > len(self.streams):
            raise RuntimeError("more lanes than streams")
        # only the streams that carry work fork from and join `main`: under
        # graph capture a wait on a stream outside the capture is invalid
        used = self.streams[:len(self.lanes)]
        for lane, s in zip(self.lanes, used):
            s.wait_stream(main)
            for n in lane:
                if n == self.SMALL:
                    _launch_small(self.small_batches, self.code, s, self.SMALL_GROUPS)
                    continue
                b = self.count[n]
                _launch(_lib.OP_ADRT, self.images[n], self.fwd[n], self.ws_f[n], self.code, b, n, s)
                _launch(_lib.OP_BDRT, self.fwd[n], self.back[n], self.ws_b[n], self.code, b, n, s)
        for s in used:
            main.wait_stream(s)

    def run(self):
        """Enqueues the whole pipeline on the current stream (eager)."""
        self._enqueue()

    def capture(self):
        """Records the pipeline as one hipGraph (after one eager warm-up)."""
        torch = _torch()
        self.run()
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._enqueue()
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            raise RuntimeError("capture() the plan before replay()")
        self.graph.replay()

    def forward(self, i):
        """Forward ADRT of image i (a view into the class buffer)."""
        n, k = self.slot[i]
        return self.fwd[n][k]

    def backprojection(self, i):
        n, k = self.slot[i]
        return self.back[n][k]

    def algorithmic_bytes(self):
        """Compulsory bytes of one pipeline run: each forward reads its image
        and writes its ADRT; each back-projection reads and writes one ADRT."""
        elem = 4 if self.code == _lib.F32 else 8
        total = 0
        for n in self.classes:
            out = 4 * n * (2 * n - 1)
            total += self.count[n] * elem * (n * n + out + 2 * out)
        return total
