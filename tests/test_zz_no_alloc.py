"""adrt_hip_adrt_noalloc: the forward ADRT without a workspace (its
intermediate lives in the library's static device scratch array).

Called through the C ABI and compared with adrt_hip_adrt byte for byte
(signs of zero included, NaNs by position), on calls in a row that reuse the
same scratch: a large input (2048^2 float32, one image per launch group: the
whole 68 MiB), then smaller ones in both dtypes (several images per group),
then a side served without any intermediate.  Sizes whose intermediate
exceeds the scratch are refused before any device work."""

import ctypes

import numpy as np
import pytest

from adrt_amd import _lib

gpu = pytest.mark.gpu


def _same(a, b):
    a, b = a.cpu().numpy(), b.cpu().numpy()
    nan = np.isnan(a) & np.isnan(b)
    return a.shape == b.shape and np.where(nan, 0, a).tobytes() == np.where(nan, 0, b).tobytes()


def _reference(lib, x, code, stream):
    import torch

    b, n = int(x.shape[0]), int(x.shape[-1])
    out = torch.empty((b, 4, 2 * n - 1, n), dtype=x.dtype, device=x.device)
    wb = lib.adrt_hip_workspace_bytes(_lib.OP_ADRT, code, b, n)
    ws = torch.empty(max(wb, 1), dtype=torch.uint8, device=x.device)
    st = lib.adrt_hip_adrt(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                           ctypes.c_void_p(ws.data_ptr()), wb, code, b, n, stream)
    assert st == _lib.OK
    return out


def _noalloc(lib, x, code, stream):
    import torch

    b, n = int(x.shape[0]), int(x.shape[-1])
    out = torch.full((b, 4, 2 * n - 1, n), float("nan"), dtype=x.dtype, device=x.device)
    st = lib.adrt_hip_adrt_noalloc(ctypes.c_void_p(x.data_ptr()),
                                   ctypes.c_void_p(out.data_ptr()), code, b, n, stream)
    return st, out


@gpu
def test_noalloc_matches_adrt_on_calls_in_a_row(cuda):
    import torch

    lib = _lib.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream(cuda).cuda_stream)
    g = torch.Generator(device=cuda)
    g.manual_seed(7)
    cases = [(3, 2048, torch.float32), (5, 256, torch.float32), (9, 128, torch.float64),
             (6, 32, torch.float32), (2, 1024, torch.float64)]
    for b, n, dt in cases:
        print(f"noalloc forward {b} x {n}^2 {dt}", flush=True)
        x = torch.randn((b, n, n), generator=g, device=cuda, dtype=dt)
        x.view(-1)[:3] = torch.tensor([-0.0, float("inf"), float("nan")], dtype=dt)
        code = _lib.F32 if dt == torch.float32 else _lib.F64
        st, got = _noalloc(lib, x, code, stream)
        assert st == _lib.OK, lib.adrt_hip_status_string(st)
        want = _reference(lib, x, code, stream)
        torch.cuda.synchronize()
        assert _same(got, want), (b, n, dt)


@gpu
def test_noalloc_refuses_sizes_beyond_the_scratch(cuda):
    import torch

    lib = _lib.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream(cuda).cuda_stream)
    x = torch.zeros((1, 2048, 2048), dtype=torch.float64, device=cuda)
    st, out = _noalloc(lib, x, _lib.F64, stream)
    torch.cuda.synchronize()
    assert st == _lib.ERR_SCRATCH
    assert bool(torch.isnan(out).all())  # refused before any device work
    assert lib.adrt_hip_adrt_noalloc(None, None, _lib.F32, 1, 2048, stream) == _lib.ERR_NULL
    assert lib.adrt_hip_adrt_noalloc(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(1),
                                     _lib.F32, 1, 48, stream) == _lib.ERR_SHAPE
