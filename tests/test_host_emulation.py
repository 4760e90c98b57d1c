"""Host emulation of the fused forward kernels (CPU, no GPU).

tests/host_emulation/fused_emu.cpp compiles the exact phase functions of
adrt_amd/csrc/fused_items.hpp for the host with AddressSanitizer and runs every
(block, thread) serially with exactly sized LDS and workspace buffers.  The
result must be byte-identical to the oracle (signed zeros included), and any
out-of-bounds access aborts the emulator.
"""

import os
import shutil
import subprocess

import numpy as np
import pytest

from helpers import canonical
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "host_emulation", "fused_emu.cpp")
ROCM_INCLUDE = "/opt/rocm/include"


@pytest.fixture(scope="module")
def emulator(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None or not os.path.isdir(ROCM_INCLUDE):
        pytest.skip("g++ or ROCm headers not available")
    exe = str(tmp_path_factory.mktemp("emu") / "fused_emu")
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-ffp-contract=off", "-fsanitize=address",
           "-D__HIP_PLATFORM_AMD__", f"-I{ROCM_INCLUDE}", f"-I{os.path.join(ROOT, 'include')}",
           "-x", "c++", SRC, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def run_emu(exe, x, tmp_path, slots=0, op="fwd", variant=0, group_bytes=0, rows=0):
    b, n = x.shape[0], x.shape[-1]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    x.tofile(fin)
    tag = "f32" if x.dtype == np.float32 else "f64"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    res = subprocess.run([exe, tag, str(b), str(n), str(fin), str(fout), str(slots), op,
                          str(variant), str(group_bytes), str(rows)],
                         capture_output=True, text=True, env=env)
    assert res.returncode == 0, res.stderr[-4000:]
    return np.fromfile(fout, dtype=x.dtype).reshape(b, 4, 2 * n - 1, n)


@pytest.mark.parametrize("dt", ["float32", "float64"])
@pytest.mark.parametrize("n", [64, 128, 256, 512, 1024])
def test_fused_bdrt_emulation(emulator, tmp_path, dt, n):
    """Fused back-projection (both passes) on the CPU: equal to the oracle
    after canonicalising the sign of zero, including Inf/NaN propagation."""
    rng = np.random.default_rng(n + 1)
    y = (rng.standard_normal((1, 4, 2 * n - 1, n)) * 10).astype(dt)
    y[0, 0, 0, :4] = [-0.0, np.inf, -np.inf, np.nan]
    y[0, 1, -1, :3] = [-0.0, -0.0, 0.0]
    got = run_emu(emulator, y, tmp_path, op="bdrt")
    assert np.array_equal(canonical(got).view(np.uint8), canonical(O.bdrt(y)).view(np.uint8))


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("dt,n,rows", [("float32", 64, 64), ("float64", 128, 1),
                                       ("float32", 256, 192), ("float32", 256, 256),
                                       ("float32", 512, 193), ("float64", 512, 384),
                                       ("float64", 512, 512), ("float32", 1024, 1024),
                                       ("float32", 1024, 1152), ("float32", 1024, 2047)])
def test_fused_bdrt_rows_emulation(emulator, tmp_path, dt, n, rows, variant):
    """Back-projection of output rows [0, rows) only (adrt_hip_bdrt_rows):
    pass B stores just those row chunks and pass A skips the items whose Z
    rows the kept rows do not depend on (from rows + 2^kB - 1 on).  Z starts
    as NaN, and after pass A every Z row from that bound on is set back to
    NaN; the data is finite, so a kept row that depends on such a row comes
    out NaN (one row less in the bound fails here).  Kept rows must equal
    the oracle's, and the rows from `rows` on stay unwritten."""
    rng = np.random.default_rng(7 * n + rows)
    y = (rng.standard_normal((2, 4, 2 * n - 1, n)) * 10).astype(dt)
    got = run_emu(emulator, y, tmp_path, op="bdrt", variant=variant, rows=rows)
    want = O.bdrt(y)
    assert not np.isnan(got[:, :, :rows]).any()
    assert np.array_equal(canonical(got[:, :, :rows]).view(np.uint8),
                          canonical(want[:, :, :rows]).view(np.uint8))
    assert np.isnan(got[:, :, rows:]).all()


@pytest.mark.parametrize("slots", [8, 24])
def test_fused_bdrt_emulation_capped_grid(emulator, tmp_path, slots):
    y = np.random.default_rng(slots).standard_normal((3, 4, 127, 64))
    got = run_emu(emulator, y, tmp_path, slots, op="bdrt")
    assert np.array_equal(canonical(got).view(np.uint8), canonical(O.bdrt(y)).view(np.uint8))


@pytest.mark.parametrize("dt", ["float32", "float64"])
@pytest.mark.parametrize("n", [64, 128, 256, 512])
def test_fused_emulation_bytes_equal(emulator, tmp_path, dt, n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal((2, n, n)) * 10).astype(dt)
    x[0, 0, :4] = [-0.0, 0.0, np.inf, np.nan]
    x[1, -1, -2:] = [1e-40 if dt == "float32" else 1e-310, -0.0]
    got = run_emu(emulator, x, tmp_path)
    want = O.adrt(x)
    same = np.isnan(got) & np.isnan(want)
    assert np.array_equal(np.where(same, 0, got).view(np.uint8), np.where(same, 0, want).view(np.uint8))


def test_fused_emulation_multi_group(emulator, tmp_path):
    """2048^2 float32: one image per group, so two images take two groups."""
    x = np.random.default_rng(7).random((2, 2048, 2048), dtype=np.float32)
    got = run_emu(emulator, x, tmp_path)
    assert np.array_equal(got.view(np.uint8), O.adrt(x).view(np.uint8))


@pytest.mark.parametrize("slots", [8, 24])
def test_fused_emulation_capped_grid(emulator, tmp_path, slots):
    """A grid smaller than the item count: each work-group walks several
    items and every item must still be computed exactly once."""
    x = np.random.default_rng(slots).standard_normal((3, 256, 256)).astype(np.float32)
    got = run_emu(emulator, x, tmp_path, slots)
    assert np.array_equal(got.view(np.uint8), O.adrt(x).view(np.uint8))


@pytest.mark.parametrize("dt", ["float32", "float64"])
@pytest.mark.parametrize("n", [128, 512, 2048])
def test_fused_alt_split_emulation(emulator, tmp_path, n, dt):
    """The other stage split (ALT_SPLIT) that the autotuner may pick for odd
    log2 n: forward pass 1 takes ceil(L/2) stages where the default takes
    floor(L/2) (float64, or L <= 9) and floor(L/2) where the default takes
    ceil(L/2) (float32 with L > 9); back-projection pass A takes floor(L/2)
    for L <= 9 and ceil(L/2) otherwise.  In float64 at 2048 this is the
    160 KB-LDS pass 1."""
    rng = np.random.default_rng(n + 5)
    x = (rng.standard_normal((1, n, n)) * 10).astype(dt)
    x[0, 0, :3] = [-0.0, np.inf, np.nan]
    got = run_emu(emulator, x, tmp_path, variant=1)
    want = O.adrt(x)
    same = np.isnan(got) & np.isnan(want)
    assert np.array_equal(np.where(same, 0, got).view(np.uint8), np.where(same, 0, want).view(np.uint8))
    y = (rng.standard_normal((1, 4, 2 * n - 1, n)) * 10).astype(dt)
    got = run_emu(emulator, y, tmp_path, op="bdrt", variant=1)
    assert np.array_equal(canonical(got).view(np.uint8), canonical(O.bdrt(y)).view(np.uint8))


@pytest.mark.parametrize("op", ["fwd", "bdrt"])
def test_fused_small_groups_emulation(emulator, tmp_path, op):
    """A group budget of two images' intermediate: five images run as
    groups of 2, 2 and 1 (the tuner's smaller launch groups)."""
    n = 128
    rng = np.random.default_rng(9)
    if op == "fwd":
        x = rng.standard_normal((5, n, n)).astype(np.float32)
        per = 4 * n * 224 * 4  # X bytes per image at n = 128 (row pitch 224)
        got = run_emu(emulator, x, tmp_path, op=op, variant=1, group_bytes=2 * per)
        assert np.array_equal(got.view(np.uint8), O.adrt(x).view(np.uint8))
    else:
        y = rng.standard_normal((5, 4, 2 * n - 1, n)).astype(np.float32)
        per = 4 * n * 320 * 4  # Z bytes per image at n = 128 (row pitch 320)
        got = run_emu(emulator, y, tmp_path, op=op, group_bytes=2 * per)
        assert np.array_equal(canonical(got).view(np.uint8), canonical(O.bdrt(y)).view(np.uint8))


IADRT_SRC = os.path.join(HERE, "host_emulation", "iadrt_emu.cpp")


@pytest.fixture(scope="module")
def iadrt_emulator(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None or not os.path.isdir(ROCM_INCLUDE):
        pytest.skip("g++ or ROCm headers not available")
    exe = str(tmp_path_factory.mktemp("emu") / "iadrt_emu")
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-ffp-contract=off", "-fsanitize=address",
           "-D__HIP_PLATFORM_AMD__", f"-I{ROCM_INCLUDE}", f"-I{os.path.join(ROOT, 'include')}",
           "-x", "c++", IADRT_SRC, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


@pytest.mark.parametrize("dt", ["float32", "float64"])
@pytest.mark.parametrize("n", [1, 2, 4, 16, 64, 128, 256])
def test_iadrt_emulation_bytes_equal(iadrt_emulator, tmp_path, dt, n):
    """The iadrt stage kernel's lane logic (lagged odd chains, 64-row ring,
    row flushes) run on the CPU must equal the oracle byte for byte."""
    rng = np.random.default_rng(n)
    b = 2 if n <= 16 else 1
    x = (rng.standard_normal((b, 4, 2 * n - 1, n)) * 10).astype(dt)
    x.flat[:3] = [-0.0, 0.0, -0.0]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    x.tofile(fin)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    res = subprocess.run([iadrt_emulator, "f32" if dt == "float32" else "f64", str(b), str(n),
                          str(fin), str(fout)], capture_output=True, text=True, env=env)
    assert res.returncode == 0, res.stderr[-4000:]
    got = np.fromfile(fout, dtype=dt).reshape(x.shape)
    assert np.array_equal(got.view(np.uint8), O.iadrt(x).view(np.uint8))


@pytest.mark.parametrize("dt", ["float32", "float64"])
@pytest.mark.parametrize("n,tail", [(4, 2), (8, 3), (32, 5), (64, 2), (64, 4), (64, 6),
                                    (128, 3), (128, 6), (256, 5), (256, 6), (256, 0),
                                    (1024, 5), (1024, 0)])
def test_iadrt_fused_tail_emulation(iadrt_emulator, tmp_path, dt, n, tail):
    """The fused tail (the last `tail` stages in one pass: lagged first
    stage, LDS rings between stages, each later stage one row behind) run on
    the CPU with every ring read checked against the row it must hold, equals
    the oracle byte for byte, for every tail length either dtype allows.  At
    n = 1024 (16 waves per plane) the stages before the tail run as
    two-stage passes (or, with no tail, as two-stage passes and one paired
    stage)."""
    if n == 1024 and dt == "float64":
        pytest.skip("one dtype suffices at this size (float64 two-stage passes: 128 and 512 below)")
    rng = np.random.default_rng(n + tail)
    b = 2 if n <= 16 else 1
    x = (rng.standard_normal((b, 4, 2 * n - 1, n)) * 10).astype(dt)
    x.flat[:3] = [-0.0, 0.0, -0.0]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    x.tofile(fin)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    res = subprocess.run([iadrt_emulator, "f32" if dt == "float32" else "f64", str(b), str(n),
                          str(fin), str(fout), str(tail)], capture_output=True, text=True, env=env)
    assert res.returncode == 0, res.stderr[-4000:]
    got = np.fromfile(fout, dtype=dt).reshape(x.shape)
    assert np.array_equal(got.view(np.uint8), O.iadrt(x).view(np.uint8))


@pytest.mark.parametrize("dt,n,tail,quad", [("float32", 128, 5, 1), ("float64", 128, 5, 1),
                                           ("float32", 512, 5, 1), ("float64", 512, 2, 1),
                                           ("float32", 512, 5, 0), ("float32", 2048, 5, 1),
                                           ("float32", 64, 0, 1), ("float64", 32, 0, 1)])
def test_iadrt_two_stage_pass_emulation(iadrt_emulator, tmp_path, dt, n, tail, quad):
    """Two paired stages per pass (the quad kernel: the four stage-(k+1)
    chain types fed by lane pairs, the B / C / D rings with every flush
    checked against the row its slot holds) and, with quad 0, one paired stage
    per pass: byte-equal to the oracle (signs of zero included), at sizes
    where the passes run 1, 2 and 3 of them before the tail (512: 2 plus one
    single, 2048: 3) and with no tail (Co of the second stage 16)."""
    rng = np.random.default_rng(n + 3 * tail + quad)
    b = 1
    x = (rng.standard_normal((b, 4, 2 * n - 1, n)) * 10).astype(dt)
    x.flat[:3] = [-0.0, 0.0, -0.0]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    x.tofile(fin)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    res = subprocess.run([iadrt_emulator, "f32" if dt == "float32" else "f64", str(b), str(n),
                          str(fin), str(fout), str(tail), str(quad)], capture_output=True,
                         text=True, env=env)
    assert res.returncode == 0, res.stderr[-4000:]
    got = np.fromfile(fout, dtype=dt).reshape(x.shape)
    assert np.array_equal(got.view(np.uint8), O.iadrt(x).view(np.uint8))


SMALL_SRC = os.path.join(HERE, "host_emulation", "small_emu.cpp")


@pytest.fixture(scope="module")
def small_emulator(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None or not os.path.isdir(ROCM_INCLUDE):
        pytest.skip("g++ or ROCm headers not available")
    exe = str(tmp_path_factory.mktemp("emu") / "small_emu")
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-ffp-contract=off", "-fsanitize=address",
           "-D__HIP_PLATFORM_AMD__", f"-I{ROCM_INCLUDE}", f"-I{os.path.join(ROOT, 'include')}",
           "-x", "c++", SMALL_SRC, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def run_small(exe, arrays, tmp_path, mode="fused", grid=64):
    """Runs the one-launch kernel's schedule over `arrays` (one per class:
    images (B, n, n), or ADRT data for mode 'bdrt'); returns per class
    (adrt, bdrt) with None for what the mode does not write."""
    args = []
    for i, a in enumerate(arrays):
        n = a.shape[-1]
        fin = tmp_path / f"in{i}.bin"
        a.tofile(fin)
        args += [str(n), str(a.shape[0]), str(fin), str(tmp_path / f"f{i}.bin"),
                 str(tmp_path / f"b{i}.bin")]
    tag = "f32" if arrays[0].dtype == np.float32 else "f64"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    res = subprocess.run([exe, tag, mode, str(grid), *args], capture_output=True, text=True,
                         env=env)
    assert res.returncode == 0, res.stderr[-4000:]
    out = []
    for i, a in enumerate(arrays):
        b, n = a.shape[0], a.shape[-1]
        shape = (b, 4, 2 * n - 1, n)
        f = None if mode == "bdrt" else np.fromfile(tmp_path / f"f{i}.bin", a.dtype).reshape(shape)
        k = None if mode == "fwd" else np.fromfile(tmp_path / f"b{i}.bin", a.dtype).reshape(shape)
        out.append((f, k))
    return out


def _same(a, b):
    """Bitwise equal, signs of zero included; NaNs match by position only
    (their payload and sign depend on the order of two NaN operands)."""
    na, nb = np.isnan(a), np.isnan(b)
    return (a.shape == b.shape and np.array_equal(na, nb)
            and np.array_equal(np.where(na, 0, a).view(np.uint8), np.where(nb, 0, b).view(np.uint8)))


@pytest.mark.parametrize("dt", ["float32", "float64"])
@pytest.mark.parametrize("grid", [32, 64, 512])
def test_small_kernel_emulation_classes(small_emulator, tmp_path, dt, grid):
    """Every side 1..64 in one persistent launch (forward, then the
    back-projection of the forward result from LDS): bit-identical to the
    oracle's adrt and bdrt(adrt), signs of zero included, NaN where the
    oracle has NaN."""
    rng = np.random.default_rng(grid)
    sides = [64, 32, 16, 8, 4, 2, 1]
    counts = [3, 5, 2, 7, 1, 3, 2]
    arrays = [(rng.standard_normal((b, n, n)) * 4).astype(dt) for n, b in zip(sides, counts)]
    arrays[0][0, 0, :4] = [-0.0, np.inf, -np.inf, np.nan]
    arrays[1][1, :, 3] = -0.0
    got = run_small(small_emulator, arrays, tmp_path, grid=grid)
    for a, (f, k) in zip(arrays, got):
        want = O.adrt(a)
        assert _same(f, want), a.shape
        assert _same(k, O.bdrt(want)), a.shape


@pytest.mark.parametrize("mode", ["fwd", "bdrt"])
@pytest.mark.parametrize("n", [64, 32, 8])
def test_small_kernel_emulation_single_op(small_emulator, tmp_path, mode, n):
    """One op over one batch, as adrt_hip_adrt / adrt_hip_bdrt run it below
    the fused range (a batch that leaves the last chunk partly empty)."""
    rng = np.random.default_rng(n)
    b = 5
    if mode == "fwd":
        x = rng.standard_normal((b, n, n)).astype(np.float32)
        (f, _), = run_small(small_emulator, [x], tmp_path, mode="fwd", grid=32)
        assert _same(f, O.adrt(x))
    else:
        y = rng.standard_normal((b, 4, 2 * n - 1, n))
        y[0, 2, -1, :2] = [-0.0, np.nan]
        (_, k), = run_small(small_emulator, [y], tmp_path, mode="bdrt", grid=32)
        assert _same(k, O.bdrt(y))
