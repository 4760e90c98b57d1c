"""Benchmark: forward ADRT of 16 ellipse phantoms at 2048 x 2048, float32.

A step is one forward ADRT (adrt_hip_adrt through the C ABI) over the whole
batch, with the input already resident in HBM.  K timed steps are bracketed
by torch.cuda.synchronize(); every call is also bracketed by HIP events on
the stream the kernels run on (torch's current stream), which gives the
per-call device time used for the roofline.

Roofline basis (compulsory traffic, each input read once and each output
written once): B_fwd(N) = s * (N^2 + 4 N (2N - 1)) bytes per image,
B_bdrt(N) = B_iadrt(N) = 2 s 4N(2N - 1); achieved = batch * B / (device
time per call); peak = 8000 GB/s.

Besides the headline line's own fields, the default run reports (key
"legs"), each with images/s, algorithmic GB/s, the fraction of peak, a
byte check of image 0 against the oracle in the same dtype and the error of
image 0 against an FP64 oracle run on the same input ("fp64_error" =
max |result - FP64| / max |FP64|):
  * bdrt and iadrt of the forward's output (the rest of the 2048^2 round
    trip), then forward, bdrt and iadrt of the same phantoms in the other
    dtype;
  * the uneven-size batch (1000 float32 images, sides 32..1024 with
    p ~ 1/n, forward + bdrt, autotuned, one hipGraph replay);
  * least squares through autograd (32 x 512^2 float32, 200 CGLS steps),
    with the float64 run of the same iteration as its FP64 reference;
and (key "cg_time_to_1e-4") the learned-vs-plain CG time to relative
residual 1e-4 on 1000 held-out 256^2 systems, 5 interleaved rounds, min /
median / max per method.  The iadrt's fp64_error in float32 is large by the
nature of the exact inverse (it amplifies the input's rounding error; the
reference's own float32 inverse behaves the same way), while its bytes
equal the oracle's.

The CPU baseline times the reference's own OpenMP CPU path (compiled from
the reference sources into oracle/_ref/libadrt_ref.so) on a bounded sample
(at least 4) of the same phantoms, on this host's cores; if that library is
absent the oracle restatement is timed instead.  The CPU model is reported.

Measured HBM traffic ("roofline.traffic") comes from the committed rocprofv3
summary profiles/fwd_adrt_16x2048_f32.md, made by two counter runs of
`bench.py --profile` (FETCH_SIZE, then WRITE_SIZE) and `bench.py --summarize`.
Each counter is scaled by what it reads for a plain 256 MiB device copy in
the same run (FETCH_SIZE reads half the bytes a copy moves, WRITE_SIZE all).

Prints exactly one JSON line (or, with --summarize, the profile summary).
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_GBS = 8000.0
METRIC = "fwd ADRT 2048² fp32 imgs/s (%HBM peak)"
PROFILE = os.path.join(ROOT, "profiles", "fwd_adrt_16x2048_f32.md")
CALIB_BYTES = 256 << 20  # device-to-device copy timed and counted beside the forward


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--batch", type=int, default=16)
    p.add_argument("--n", type=int, default=2048)
    p.add_argument("--dtype", default="float32", choices=["float32", "float64"])
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--no-legs", action="store_true",
                   help="time only the headline forward (skip bdrt / iadrt / the other "
                        "dtype / uneven batch / least squares)")
    p.add_argument("--no-learned", action="store_true",
                   help="skip the learned-vs-plain CG time-to-1e-4 leg")
    p.add_argument("--cg-count", type=int, default=1000)
    p.add_argument("--cg-repeats", type=int, default=5,
                   help="interleaved rounds of the plain / learned / preconditioner-only solves")
    p.add_argument("--profile", action="store_true",
                   help="short run for rocprofv3 (1 warm-up + 3 steps, no CPU leg), followed "
                        "by a 256 MiB device copy that calibrates the memory counters")
    p.add_argument("--summarize", nargs="+", metavar="DIR",
                   help="print a summary of rocprofv3 csv output directories (kernel stats "
                        "and/or FETCH_SIZE / WRITE_SIZE counter runs of --profile)")
    p.add_argument("--calls", type=int, default=4,
                   help="calls in the profiled run (--summarize)")
    p.add_argument("--match", default="adrt_hip",
                   help="count only kernels whose name contains this (--summarize); with "
                        "--iterations, the kernel that ends each iteration")
    p.add_argument("--iterations", type=int, default=0,
                   help="--summarize a kernel trace per solver iteration over the last K "
                        "iterations (e.g. --op pcg: --match pcg_dir_kernel --iterations 5)")
    p.add_argument("--sequence", type=int, default=0,
                   help="--summarize a kernel trace as the K dispatches that start at the "
                        "last one whose name contains --match (e.g. --op unet: --match "
                        "encoder0 --sequence 16, one row per layer)")
    p.add_argument("--op", default="adrt",
                   choices=["adrt", "bdrt", "iadrt", "unet", "pcg", "uneven", "lsq"],
                   help="operation a --profile run calls (bdrt/iadrt on the forward output)")
    return p.parse_args()


def bytes_fwd(n, s):
    return s * (n * n + 4 * n * (2 * n - 1))


def bytes_bdrt(n, s):
    return 2 * s * 4 * n * (2 * n - 1)


def time_calls(fn, steps, warmup, stream):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record(stream)
        fn()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dev_ms = sum(s.elapsed_time(e) for s, e in zip(starts, ends)) / steps
    return wall, dev_ms


UNEVEN_SIDES = (32, 64, 128, 256, 512, 1024)
LEARNED_WEIGHTS = os.path.join(ROOT, "adrt_amd", "weights", "learned_256.npz")


def lsq_leg(iters=200):
    """C4: least squares min 0.5 ||adrt(x) - b||^2 through autograd for 32
    float32 images of 512x512 from x = 0, b = adrt(phantom) + N(0, (0.01
    max|b|)^2), ``iters`` CGLS steps whose A^T is the autograd backward
    (the fused back-projection).  Each step applies A once and A^T once."""
    import torch

    from adrt_amd import _lib, autograd, batched, learned

    # the autotuner's kernel configuration for this batch shape
    tuning = [batched.tune(op, 512, 32)["choice"] for op in (_lib.OP_ADRT, _lib.OP_BDRT)]
    x_true = learned.phantom_batch(512, 4, 0, 32)
    a = learned._kernels.adrt(x_true)
    g = torch.Generator(device=a.device)
    g.manual_seed(4)
    b = a + 0.01 * a.abs().amax() * torch.randn(a.shape, generator=g, device=a.device)
    autograd.lsq_cg(b, iters=2)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, hist = autograd.lsq_cg(b, iters=iters)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    per_iter = bytes_fwd(512, 4) * 32 * 2  # A and A^T, compulsory bytes each
    gbs = per_iter / (secs / iters) / 1e9
    # the same iteration in float64 (untimed): the error of the float32 result
    # against an FP64 run
    x64, _ = autograd.lsq_cg(b.double(), iters=iters)
    return {
        "workload": f"least squares through autograd (CG on the normal equations, hipGraph "
                    f"blocks), 32 x 512x512 float32, {iters} steps, x0 = 0",
        "tuned_adrt_bdrt": tuning,
        "seconds": round(secs, 4),
        "ms_per_step": round(secs / iters * 1e3, 3),
        "loss_first": float(hist[0].mean()), "loss_last": float(hist[-1].mean()),
        "rel_error_vs_phantom": float(((x - x_true).norm() / x_true.norm())),
        "GBps_algorithmic": round(gbs, 1),
        "frac": round(gbs / PEAK_GBS, 4),
        "fp64_error": float((x.double() - x64).abs().max() / x64.abs().max()),
    }


def learned_cg_leg(count, repeats):
    """Second half of the metric: wall time to relative residual 1e-4 of CG
    on the normal equations for ``count`` held-out 256x256 float32 phantoms
    (noise 0.5% of max|adrt(x)|), learned (U-Net initial guess + learned
    multiscale preconditioner) vs plain CG, on the same GPU in the same run,
    ``repeats`` interleaved rounds of (plain, learned, preconditioner only).
    Uses the committed weights trained on the GPU by
    ``python -m adrt_amd.train_learned``; trains for 60 s if they are absent."""
    from adrt_amd import learned

    if os.path.exists(LEARNED_WEIGHTS):
        solver = learned.LearnedSolver.load(LEARNED_WEIGHTS)
        source = "adrt_amd/weights/learned_256.npz"
    else:
        solver = learned.LearnedSolver(n=256)
        solver.fit_preconditioner()
        solver.train(2000, batch=64)
        source = "trained in this run"
    out = learned.evaluate(solver, count=count, repeats=repeats)
    out["weights"] = source
    out["workload"] = (f"{count} held-out 256x256 float32 ellipse phantoms (seed 2000), "
                       "d = adrt(x) + 0.5% noise, CG on mean_q truncate(bdrt(adrt(.)))")
    return out


def uneven_extra(steps, warmup):
    """1000 float32 images, sides drawn from UNEVEN_SIDES with probability
    proportional to 1/n (numpy default_rng(0)), pixels U[0, 1) (torch
    generator seed 0): forward ADRT + back-projection of all of them,
    autotuned per side and between the one-launch queue and the per-side
    calls on streams, captured as one hipGraph; K graph replays timed."""
    import numpy as np
    import torch

    from adrt_amd import batched

    sides = np.array(UNEVEN_SIDES)
    p = 1.0 / sides
    sizes = np.random.default_rng(0).choice(sides, 1000, p=p / p.sum()).tolist()
    t0 = time.perf_counter()
    plan = batched.UnevenBatch(sizes, cache=batched.TuneCache())
    tune_s = time.perf_counter() - t0
    g = torch.Generator(device=plan.device)
    g.manual_seed(0)
    for n in plan.classes:
        plan.images[n].uniform_(0.0, 1.0, generator=g)
    plan.capture()
    stream = torch.cuda.current_stream(plan.device)
    wall, graph_ms = time_calls(plan.replay, steps, warmup, stream)
    _, eager_ms = time_calls(plan.run, steps, warmup, stream)
    algo = plan.algorithmic_bytes()
    gbs = algo / (graph_ms * 1e-3) / 1e9
    choice = {f"{t['op']}{t['n']}": [t["choice"]["variant"], t["choice"]["group_bytes"] >> 20]
              for t in plan.tuning}
    pipeline = ("queue (one persistent launch, bundle %d, %d waves)"
                % (plan.queue.bundle, plan.queue.waves) if plan.queue is not None
                else "streams (a forward and a back-projection call per side)")
    timeouts = plan.queue.timeouts() if plan.queue is not None else 0
    # checks on the first image of every side: the replayed outputs against
    # the oracle (bytes; bdrt up to the sign of zero) and against an FP64
    # oracle run on the same image
    from oracle import oracle as O

    exact, err = True, 0.0
    for n in plan.classes:
        i = plan.slot.index((n, 0))
        x = plan.images[n][:1].cpu().numpy()
        y, z = plan.forward(i).cpu().numpy()[None], plan.backprojection(i).cpu().numpy()[None]
        y_ref = O.adrt(x)
        exact = exact and same_bytes(y, y_ref) and same_bytes_up_to_zero_sign(z, O.bdrt(y_ref))
        y64 = O.adrt(x.astype(np.float64))
        err = max(err, fp64_error(y, y64), fp64_error(z, O.bdrt(y.astype(np.float64))))
    return {
        "workload": "forward ADRT + bdrt of 1000 float32 images, n in {32..1024} with "
                    "p ~ 1/n (seed 0), one hipGraph",
        "counts": {str(n): plan.count[n] for n in plan.classes},
        "graph_ms_per_pipeline": round(graph_ms, 3),
        "eager_ms_per_pipeline": round(eager_ms, 3),
        "imgs_per_s": round(1000 / graph_ms * 1e3, 1),
        "algorithmic_bytes": algo,
        "GBps_algorithmic": round(gbs, 1),
        "frac": round(gbs / PEAK_GBS, 4),
        "bit_exact_first_image_per_side": exact and timeouts == 0,
        "fp64_error": err,
        "autotune_s": round(tune_s, 2),
        "pipeline": pipeline,
        "tune_cache": {"path": batched.TuneCache().path,
                       "hit": plan.queue_cached and all(x["cached"] for x in plan.tuning)},
        "tuned_ms": {"streams": plan.streams_ms and round(plan.streams_ms, 3),
                     "queue": [[q["bundle"], q["waves"], round(q["ms"], 3)]
                               for q in plan.queue_timings]},
        "tuned_variant_groupMiB": choice,
    }


def cpu_model():
    """The host CPU's model name (first "model name" line of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(host_imgs, budget_s):
    """Time the reference's CPU path on a bounded sample; returns a dict."""
    import numpy as np

    from oracle import oracle as O
    from oracle import ref as R

    kind = "reference" if R.available() else "port"
    fn = R.adrt if kind == "reference" else O.adrt
    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    one = np.ascontiguousarray(host_imgs[:1])
    fn(one)  # warm-up (first touch of the library and the thread pool)
    t0 = time.perf_counter()
    fn(one)
    est = time.perf_counter() - t0
    runs = 5
    reps = max(min(4, len(host_imgs)),
               min(len(host_imgs), int(budget_s / max(est, 1e-3) / runs)))
    sample = np.ascontiguousarray(host_imgs[:reps])
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn(sample)
        times.append(time.perf_counter() - t0)
    med = sorted(times)[runs // 2]
    n = host_imgs.shape[-1]
    out = {
        "value": reps / med,
        "unit": "imgs/s",
        "cores": cores,
        "kind": kind,
        "cpu_model": cpu_model(),
        "sample": f"adrt of {reps} of the same {n}x{n} {host_imgs.dtype} phantoms, "
                  f"median of {runs} runs after a warm-up, OpenMP threads={cores}",
        "seconds_per_run_median": round(med, 4),
    }
    if kind == "reference":
        # the same call on one thread (the reference's OpenMP runtime is
        # libgomp, shared with this process): one image, median of 3
        import ctypes

        try:
            gomp = ctypes.CDLL("libgomp.so.1")
            gomp.omp_set_num_threads(1)
            t1 = []
            for _ in range(3):
                t0 = time.perf_counter()
                fn(one)
                t1.append(time.perf_counter() - t0)
            gomp.omp_set_num_threads(cores)
            out["single_thread_imgs_per_s"] = 1.0 / sorted(t1)[1]
        except OSError:
            pass
    return out


def _csv_rows(dirs, suffix):
    import csv
    import glob

    for d in dirs:
        for f in sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True)):
            with open(f, newline="") as fh:
                yield from csv.DictReader(fh)


def _short(name):
    name = name.split("(")[0].replace("void ", "")
    return name.replace("adrt_hip::fused::", "")


def summarize(dirs, calls, match="adrt_hip"):
    """Kernel durations and calibrated HBM traffic per forward call from
    rocprofv3 csv output (kernel stats and FETCH_SIZE / WRITE_SIZE runs)."""
    lines = []
    stats = [r for r in _csv_rows(dirs, "kernel_stats.csv") if match in r["Name"]]
    if stats:
        lines += ["| kernel | calls | avg µs | min µs | max µs | total µs |",
                  "|---|---|---|---|---|---|"]
        total = 0.0
        for r in stats:
            t = float(r["TotalDurationNs"]) / 1e3
            total += t
            lines.append(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                         f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | {t:.1f} |")
        lines.append(f"\nkernel_us_per_call: {total / calls:.1f}")
    ctr = {}
    for r in _csv_rows(dirs, "counter_collection.csv"):
        name = r.get("Kernel_Name", "")
        key = ("copy" if "copyBuffer" in name else _short(name)) if (
            match in name or "copyBuffer" in name) else None
        if key is None:
            continue
        ctr.setdefault(key, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    if ctr:
        copy = ctr.get("copy", {})
        kib = float(1 << 10)
        # the calibration copy is the largest copy dispatch of the run
        fetch_scale = CALIB_BYTES / (max(copy["FETCH_SIZE"]) * kib) if copy.get("FETCH_SIZE") else None
        write_scale = CALIB_BYTES / (max(copy["WRITE_SIZE"]) * kib) if copy.get("WRITE_SIZE") else None
        lines += ["", f"calibration (256 MiB device copy): FETCH_SIZE x {fetch_scale:.3f}, "
                      f"WRITE_SIZE x {write_scale:.3f}" if fetch_scale and write_scale else
                  "calibration copy missing", "",
                  "| kernel | dispatches | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch | "
                  "HBM read MiB/dispatch | HBM write MiB/dispatch |", "|---|---|---|---|---|---|"]
        per_call = 0.0
        for key, d in sorted(ctr.items()):
            if key == "copy":
                continue
            f = d.get("FETCH_SIZE", [])
            w = d.get("WRITE_SIZE", [])
            fm = sum(f) / len(f) if f else float("nan")
            wm = sum(w) / len(w) if w else float("nan")
            rd = fm * kib * (fetch_scale or float("nan")) / (1 << 20)
            wr = wm * kib * (write_scale or float("nan")) / (1 << 20)
            lines.append(f"| `{key}` | {max(len(f), len(w))} | {fm:.0f} | {wm:.0f} | {rd:.1f} | {wr:.1f} |")
            per_call += (sum(f) * (fetch_scale or 0) + sum(w) * (write_scale or 0)) * kib
        lines.append(f"\ntraffic_bytes_per_call: {int(per_call / calls)}")
    return "\n".join(lines)


def summarize_iterations(dirs, marker, last):
    """Per-iteration kernel time of the last ``last`` iterations of a traced
    solver run (rocprofv3 --kernel-trace csv): the dispatches after the
    ``last + 1``-th from last dispatch whose name contains ``marker`` (a
    kernel that runs once per iteration, at its end), grouped by kernel."""
    rows = sorted(_csv_rows(dirs, "kernel_trace.csv"), key=lambda r: int(r["Start_Timestamp"]))
    ends = [i for i, r in enumerate(rows) if marker in r["Kernel_Name"]]
    if len(ends) < last + 1:
        return f"fewer than {last + 1} dispatches of {marker}"
    window = rows[ends[-last - 1] + 1:ends[-1] + 1]
    agg = {}
    for r in window:
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        a = agg.setdefault(_short(r["Kernel_Name"]), [0, 0.0])
        a[0] += 1
        a[1] += d
    lines = ["| kernel | dispatches / iteration | µs / iteration |", "|---|---|---|"]
    total = 0.0
    for k, (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        total += t
        lines.append(f"| `{k}` | {c / last:.1f} | {t / last:.1f} |")
    lines.append(f"\nkernel_us_per_iteration: {total / last:.1f}")
    return "\n".join(lines)


def summarize_sequence(dirs, marker, count):
    """The ``count`` dispatches of a traced run (rocprofv3 --kernel-trace csv)
    that start at the last one whose name contains ``marker``, in order,
    with their durations."""
    rows = sorted(_csv_rows(dirs, "kernel_trace.csv"), key=lambda r: int(r["Start_Timestamp"]))
    starts = [i for i, r in enumerate(rows) if marker in r["Kernel_Name"]]
    if not starts:
        return f"no dispatch of {marker}"
    lines = ["| # | kernel | µs |", "|---|---|---|"]
    total = 0.0
    for k, r in enumerate(rows[starts[-1]:starts[-1] + count]):
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        total += d
        lines.append(f"| {k} | `{_short(r['Kernel_Name'])}` | {d:.1f} |")
    lines.append(f"\nkernel_us_total: {total:.1f}")
    return "\n".join(lines)


def profiled_traffic():
    """HBM bytes per forward call from the committed rocprofv3 summary."""
    try:
        with open(PROFILE) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_call:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def same_bytes(got, want):
    """Raw byte equality (signs of zero included), NaNs matched by position."""
    import numpy as np

    got, want = np.asarray(got), np.asarray(want)
    if got.shape != want.shape or got.dtype != want.dtype:
        return False
    nan = np.isnan(got) & np.isnan(want)
    return bool(np.where(nan, 0, got).tobytes() == np.where(nan, 0, want).tobytes())


def same_bytes_up_to_zero_sign(got, want):
    import numpy as np

    got = np.where(np.asarray(got) == 0, 0, got).astype(np.asarray(got).dtype)
    want = np.where(np.asarray(want) == 0, 0, want).astype(np.asarray(want).dtype)
    return same_bytes(got, want)


def fp64_error(got, ref64):
    """max |got - ref64| / max |ref64| over the compared entries: the error of
    a result against an FP64 run of the oracle on the same input."""
    import numpy as np

    ref64 = np.asarray(ref64, dtype=np.float64)
    scale = float(np.abs(ref64).max())
    err = float(np.abs(np.asarray(got, dtype=np.float64) - ref64).max())
    return err / scale if scale > 0 else err


def transform_leg(lib, name, x, steps, stream):
    """Times adrt_hip_<name> (adrt, bdrt or iadrt) on the device batch x
    through the C ABI (HIP events on the kernels' stream, one warm-up call),
    then checks image 0: bytes against the oracle in x's dtype (bdrt: up to
    the sign of zero, see the fused back-projection's boundary rule) and the
    error against an FP64 oracle run on the same input.  Returns (output,
    record)."""
    import ctypes

    import numpy as np
    import torch

    from adrt_amd import _lib
    from oracle import oracle as O

    op = {"adrt": _lib.OP_ADRT, "bdrt": _lib.OP_BDRT, "iadrt": _lib.OP_IADRT}[name]
    code = _lib.F32 if x.dtype == torch.float32 else _lib.F64
    batch, n, s = int(x.shape[0]), int(x.shape[-1]), x.element_size()
    shape = (batch, 4, 2 * n - 1, n) if name == "adrt" else tuple(x.shape)
    out = torch.empty(shape, device=x.device, dtype=x.dtype)
    wb = lib.adrt_hip_workspace_bytes(op, code, batch, n)
    ws = torch.empty(max(wb, 1), device=x.device, dtype=torch.uint8)
    fn_c = getattr(lib, f"adrt_hip_{name}")
    call = (ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
            ctypes.c_void_p(ws.data_ptr()), wb, code, batch, n,
            ctypes.c_void_p(stream.cuda_stream))

    def run():
        st = fn_c(*call)
        if st != 0:
            _lib.check(st, f"adrt_hip_{name}")

    _, ms = time_calls(run, steps, 1, stream)
    del ws
    algo = batch * (bytes_fwd(n, s) if name == "adrt" else bytes_bdrt(n, s))
    gbs = algo / (ms * 1e-3) / 1e9
    xin = x[:1].cpu().numpy()
    got = out[:1].cpu().numpy()
    ref = getattr(O, name)
    same = ref(xin)
    rec = {
        "workload": f"{name} of {batch} x {n}x{n} {x.dtype}".replace("torch.", ""),
        "imgs_per_s": round(batch / ms * 1e3, 1),
        "ms_per_call": round(ms, 4),
        "GBps_algorithmic": round(gbs, 1),
        "frac": round(gbs / PEAK_GBS, 4),
        "algorithmic_bytes_per_call": algo,
        "bit_exact_first_image": (same_bytes_up_to_zero_sign(got, same) if name == "bdrt"
                                  else same_bytes(got, same)),
        "fp64_error": fp64_error(got, same if x.dtype == torch.float64
                                 else ref(xin.astype(np.float64))),
    }
    if name == "bdrt":
        rec["bit_exact_first_image_signed_zeros"] = same_bytes(got, same)
    return out, rec


def main():
    args = parse()
    if args.summarize:
        if args.iterations:
            print(summarize_iterations(args.summarize, args.match, args.iterations))
        elif args.sequence:
            print(summarize_sequence(args.summarize, args.match, args.sequence))
        else:
            print(summarize(args.summarize, args.calls, args.match))
        return
    if args.profile:
        args.steps, args.warmup, args.no_cpu, args.no_legs = 3, 1, True, True
        args.no_learned = True
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported: "
              "this benchmark runs on a single MI355X (a batch of images would shard "
              "across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)
    import ctypes

    import numpy as np
    import torch

    import adrt_amd
    from adrt_amd import _device as dev
    from adrt_amd import _lib
    from adrt_amd.phantom import phantoms

    lib = _lib.load()
    torch.cuda.init()
    device = torch.device("cuda", 0)
    if args.profile and args.op == "unet":
        # two U-Net forwards (the learned initial guess) on 1000 random
        # 256x256 images with the committed weights
        from adrt_amd import learned

        solver = learned.LearnedSolver.load(LEARNED_WEIGHTS)
        z = torch.randn(1000, 256, 256, device=device, generator=torch.Generator(
            device=device).manual_seed(0))
        for _ in range(2):
            solver.net.forward(z, train=False)
        torch.cuda.synchronize()
        print(json.dumps({"op": "unet", "images": 1000, "n": 256}), flush=True)
        return
    if args.profile and args.op == "lsq":
        # least squares through autograd: 32 x 512^2 float32, 48 steps (one
        # eager block of 16, then two graph replays)
        from adrt_amd import autograd, learned

        x_true = learned.phantom_batch(512, 4, 0, 32)
        b = learned._kernels.adrt(x_true)
        autograd.lsq_cg(b, iters=48)
        torch.cuda.synchronize()
        print(json.dumps({"op": "lsq", "images": 32, "n": 512, "steps": 48}), flush=True)
        return
    if args.profile and args.op == "uneven":
        # the uneven-size pipeline of 1000 images (autotuned, captured):
        # 3 graph replays after the capture's eager warm-up
        import numpy as np
        from adrt_amd import batched

        sides = np.array(UNEVEN_SIDES)
        p = 1.0 / sides
        sizes = np.random.default_rng(0).choice(sides, 1000, p=p / p.sum()).tolist()
        plan = batched.UnevenBatch(sizes)
        g = torch.Generator(device=plan.device)
        g.manual_seed(0)
        for n in plan.classes:
            plan.images[n].uniform_(0.0, 1.0, generator=g)
        plan.capture()
        for _ in range(3):
            plan.replay()
        torch.cuda.synchronize()
        a = torch.empty(CALIB_BYTES, device=device, dtype=torch.uint8)  # calibration copy
        torch.empty_like(a).copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"op": "uneven", "images": 1000, "lanes": [list(map(str, lane))
                                                                    for lane in plan.lanes]}),
              flush=True)
        return
    if args.profile and args.op == "pcg":
        # preconditioned CG on the 1000 held-out 256x256 systems of the
        # learned-CG leg: 2 iterations (warm-up), then 5 iterations without
        # compaction, the pattern the per-iteration kernel breakdown uses
        from adrt_amd import batched, learned

        solver = learned.LearnedSolver.load(LEARNED_WEIGHTS)
        _, b = solver.data(2000, 0, 1000, noise_seed=2000)
        for op in (_lib.OP_ADRT, _lib.OP_BDRT):
            batched.tune(op, 256, 1000)
        learned.pcg(learned.normal_op, b, prec=solver.prec, rtol=1e-4, maxiter=2)
        learned.pcg(learned.normal_op, b, prec=solver.prec, rtol=1e-4, maxiter=5, compact=0)
        torch.cuda.synchronize()
        print(json.dumps({"op": "pcg", "images": 1000, "n": 256, "iterations": 5}), flush=True)
        return
    n, batch = args.n, args.batch
    s = 4 if args.dtype == "float32" else 8
    imgs = phantoms(batch, n, seed=0, dtype=args.dtype, device=device)
    out = torch.empty((batch, 4, 2 * n - 1, n), device=device, dtype=imgs.dtype)
    code = _lib.F32 if args.dtype == "float32" else _lib.F64
    ws_bytes = lib.adrt_hip_workspace_bytes(_lib.OP_ADRT, code, batch, n)
    ws = torch.empty(max(ws_bytes, 1), device=device, dtype=torch.uint8)
    stream = torch.cuda.current_stream(device)
    sp = ctypes.c_void_p(stream.cuda_stream)
    pin, pout, pws = (ctypes.c_void_p(t.data_ptr()) for t in (imgs, out, ws))

    def fwd():
        st = lib.adrt_hip_adrt(pin, pout, pws, ws_bytes, code, batch, n, sp)
        if st != 0:
            _lib.check(st, "adrt_hip_adrt")

    if args.profile and args.op != "adrt":
        # Profile bdrt or iadrt on the forward's output: 1 warm-up + 3 calls,
        # then the calibration copy.
        fwd()
        op_id = {"bdrt": _lib.OP_BDRT, "iadrt": _lib.OP_IADRT}[args.op]
        y = out
        z = torch.empty_like(y)
        wb = lib.adrt_hip_workspace_bytes(op_id, code, batch, n)
        wst = torch.empty(max(wb, 1), device=device, dtype=torch.uint8)
        fn_c = getattr(lib, f"adrt_hip_{args.op}")
        call = (ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(z.data_ptr()),
                ctypes.c_void_p(wst.data_ptr()), wb, code, batch, n, sp)
        _, ms = time_calls(lambda: fn_c(*call), 3, 1, stream)
        a = torch.empty(CALIB_BYTES, device=device, dtype=torch.uint8)
        torch.empty_like(a).copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"op": args.op, "ms_per_call": round(ms, 4),
                          "imgs_per_s": round(batch / ms * 1e3, 2)}), flush=True)
        return

    wall, dev_ms = time_calls(fwd, args.steps, args.warmup, stream)
    value = batch * args.steps / wall
    algo = batch * bytes_fwd(n, s)
    achieved = algo / (dev_ms * 1e-3) / 1e9
    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "imgs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if s == 4 else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"adrt forward, {batch} random-ellipse phantoms {n}x{n} {args.dtype} "
                        "(seed 0), inputs resident in HBM",
            "batch": batch,
            "n": n,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_GBS, 4),
            "traffic": profiled_traffic() if (n, batch, s) == (2048, 16, 4) else None,
            "kernel": "adrt_hip_adrt: one forward call (fwd_pass1 + fwd_pass2 per group "
                      "of images)",
            "algorithmic_bytes_per_call": algo,
            "device_ms_per_call": round(dev_ms, 4),
        },
    }
    if args.profile:
        a = torch.empty(CALIB_BYTES, device=device, dtype=torch.uint8)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps(result), flush=True)
        return
    from oracle import oracle as O

    # the headline's own checks on image 0: bytes against the oracle in the
    # same dtype, and the error against an FP64 run of the oracle
    x0 = imgs[:1].cpu().numpy()
    y0 = out[:1].cpu().numpy()
    result["parity_first_image_bit_exact"] = same_bytes(y0, O.adrt(x0))
    result["fp64_error"] = fp64_error(y0, O.adrt(x0.astype(np.float64)))
    host = imgs.cpu().numpy()
    legs = {}
    if not args.no_legs:
        steps = max(3, args.steps // 4)
        # the round trip's other legs on the forward's output, then the whole
        # round trip in float64 on the same phantoms
        y = out
        del ws
        for name in ("bdrt", "iadrt"):
            _, legs[f"{name}_{args.dtype}"] = transform_leg(lib, name, y, steps, stream)
        del y, out
        torch.cuda.empty_cache()
        other = "float64" if args.dtype == "float32" else "float32"
        x64 = phantoms(batch, n, seed=0, dtype=other, device=device)
        y64, legs[f"adrt_{other}"] = transform_leg(lib, "adrt", x64, steps, stream)
        del x64
        for name in ("bdrt", "iadrt"):
            _, legs[f"{name}_{other}"] = transform_leg(lib, name, y64, steps, stream)
        del y64
        torch.cuda.empty_cache()
        legs["uneven_c3"] = uneven_extra(steps, 2)
        torch.cuda.empty_cache()
        legs["lsq_autograd_c4"] = lsq_leg()
        torch.cuda.empty_cache()
    result["legs"] = legs
    if not args.no_learned:
        result["cg_time_to_1e-4"] = learned_cg_leg(args.cg_count, args.cg_repeats)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(host, args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
