"""Forward ADRT and back-projection of many images of uneven sizes.

``UnevenBatch`` takes a list of image sides (powers of two) and holds, per
distinct side n, one stacked device buffer of the images of that side, their
forward ADRTs and back-projections, and the workspaces.  ``run()`` enqueues,
per size class, one batched forward call and one batched back-projection
call, each size class on its own stream (so small classes overlap), and
``capture()`` records the whole pipeline once as a hipGraph that
``replay()`` relaunches with no host work.

``tune()`` is the autotuner: for one (op, n, batch) it times every kernel
variant the library offers for that size (the fused two-pass kernels with
either stage split, i.e. the number of stages fused in LDS per pass, and the
streaming one-kernel-per-stage path) and every launch-group size that makes a
difference, with HIP events on data of that shape, then leaves the fastest
configuration selected in the library.  Results are bit-identical for every
configuration; only speed differs.

The configuration is process-wide per (op, n) (``adrt_hip_set_config``): a
later ``tune`` of the same side for another batch size changes what eager
calls of an earlier plan use (a captured graph keeps the kernels it
recorded).  Plans check their workspaces against the library on every eager
call and raise if the configuration changed under them.
"""

import ctypes

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


def candidates(op, n, batch, per_image_bytes):
    """Distinct (variant, group_bytes) configurations for (op, n, batch).

    The alternative stage split exists for odd log2 n in the fused range;
    group budgets are kept only when they change the images per launch group
    (``per_image_bytes``: the fused path's intermediate per image)."""
    L = int(n).bit_length() - 1
    out = [(_lib.VARIANT_STREAMING, 0)]
    if not fused_range(n):
        return out
    seen = set()
    budgets = []
    for b in GROUP_BUDGETS:
        g = min(max(b // per_image_bytes, 1), batch)
        if g not in seen:
            seen.add(g)
            budgets.append(b)
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


def tune(op, n, batch, *, dtype=None, device=None, reps=3):
    """Times every candidate configuration of (op, n, batch) and selects the
    fastest in the library.  Returns a dict with the choice and all timings
    (milliseconds per call, mean of ``reps`` after one warm-up)."""
    torch = _torch()
    dev.require_gpu()
    dtype = dtype or torch.float32
    device = device or torch.device("cuda", torch.cuda.current_device())
    code = _code(dtype)
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
    return {"op": "adrt" if op == _lib.OP_ADRT else "bdrt", "n": n, "batch": batch,
            "choice": best, "timings": timings}


class UnevenBatch:
    """Forward ADRT + back-projection plan for images of sides ``sizes``."""

    def __init__(self, sizes, *, dtype=None, device=None, autotune=True, reps=3):
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
        self.tuning = []
        if autotune:
            for n in self.classes:
                for op in (_lib.OP_ADRT, _lib.OP_BDRT):
                    self.tuning.append(tune(op, n, count[n], dtype=self.dtype,
                                            device=self.device, reps=reps))
        self.images, self.fwd, self.back, self.ws_f, self.ws_b = {}, {}, {}, {}, {}
        for n in self.classes:
            b = count[n]
            self.images[n] = torch.zeros((b, n, n), dtype=self.dtype, device=self.device)
            self.fwd[n] = torch.empty((b, 4, 2 * n - 1, n), dtype=self.dtype, device=self.device)
            self.back[n] = torch.empty_like(self.fwd[n])
            self.ws_f[n] = _workspace(_lib.OP_ADRT, self.code, b, n, self.device)
            self.ws_b[n] = _workspace(_lib.OP_BDRT, self.code, b, n, self.device)
        self.lanes = self._assign_lanes()
        self.streams = [torch.cuda.Stream(self.device) for _ in self.lanes]
        self.graph = None

    # The runtime maps streams onto 4 hardware queues (HIP's default), so
    # more than 4 streams share queues and serialise whatever lands
    # together.  The size classes are spread over at most 4 streams by
    # longest-processing-time-first: each class (costliest first) goes to the
    # stream with the least work so far.  A class's cost is its tuned
    # forward + back-projection time, or, untuned, its bytes plus a fixed
    # per-launch latency (small sides are launch-bound).
    HW_QUEUES = 4

    def _class_cost(self, n):
        ms = [t["choice"]["ms"] for t in self.tuning if t["n"] == n and "ms" in t["choice"]]
        if len(ms) == 2:
            return sum(ms)
        launches = 4 if fused_range(n) else 2 * (n.bit_length() + 2)
        return self.count[n] * n * n * 100 / 5e9 * 1e3 + launches * 0.02  # ms

    def _assign_lanes(self):
        lanes = [[] for _ in range(min(self.HW_QUEUES, len(self.classes)))]
        load = [0.0] * len(lanes)
        for n in sorted(self.classes, key=self._class_cost, reverse=True):
            i = min(range(len(lanes)), key=lambda j: load[j])
            lanes[i].append(n)
            load[i] += self._class_cost(n)
        return lanes

    def load(self, images):
        """Copies ``images`` (a sequence of (n, n) tensors in plan order)."""
        if len(images) != len(self.sizes):
            raise ValueError(f"expected {len(self.sizes)} images, got {len(images)}")
        for img, (n, k) in zip(images, self.slot):
            if tuple(img.shape) != (n, n):
                raise ValueError(f"image of shape {tuple(img.shape)} where ({n}, {n}) was planned")
            self.images[n][k].copy_(img)

    def _enqueue(self):
        torch = _torch()
        main = torch.cuda.current_stream(self.device)
        for lane, s in zip(self.lanes, self.streams):
            s.wait_stream(main)
            for n in lane:  # costliest first
                b = self.count[n]
                _launch(_lib.OP_ADRT, self.images[n], self.fwd[n], self.ws_f[n], self.code, b, n, s)
                _launch(_lib.OP_BDRT, self.fwd[n], self.back[n], self.ws_b[n], self.code, b, n, s)
        for s in self.streams:
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
