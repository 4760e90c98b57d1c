"""Device plumbing: torch for HBM allocations and streams, nothing else.

NumPy arrays (and CPU tensors) make a host round trip: copied to the current
HIP device, transformed there, copied back.  CUDA/HIP tensors stay on their
device and run on that device's current torch stream.  Without a HIP device
every compute call raises ``RuntimeError``; there is no CPU path.
"""

import ctypes
import warnings

import numpy as np

from . import _lib

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as _t

        _torch = _t
    return _torch


def is_tensor(x):
    if _torch is None:
        # Avoid importing torch just to reject a non-tensor.
        mod = type(x).__module__
        if not mod.startswith("torch"):
            return False
    return isinstance(x, torch().Tensor)


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError(
            "adrt_amd: no HIP GPU is available; the transform core runs only on "
            "the GPU (there is no CPU fallback)"
        )
    _lib.load()


_NP_TO_TORCH = {
    np.dtype(np.float32): "float32",
    np.dtype(np.float64): "float64",
    np.dtype(np.int8): "int8",
    np.dtype(np.uint8): "uint8",
    np.dtype(np.int16): "int16",
    np.dtype(np.int32): "int32",
    np.dtype(np.int64): "int64",
    np.dtype(np.bool_): "bool",
    np.dtype(np.float16): "float16",
}


def dtype_code(arr):
    """ADRT_HIP_F32 / ADRT_HIP_F64 for a float array or tensor, else None."""
    if is_tensor(arr):
        t = torch()
        if arr.dtype == t.float32:
            return _lib.F32
        if arr.dtype == t.float64:
            return _lib.F64
        return None
    if arr.dtype == np.float32:
        return _lib.F32
    if arr.dtype == np.float64:
        return _lib.F64
    return None


def dtype_name(arr):
    return str(arr.dtype)


def element_size(arr):
    return arr.element_size() if is_tensor(arr) else arr.dtype.itemsize


def to_device(arr):
    """Return (device tensor, restore) where restore(tensor) maps the result
    back to the caller's array type and location."""
    t = torch()
    if is_tensor(arr) and arr.is_cuda:
        return arr, lambda out: out
    require_gpu()
    dev = t.device("cuda", t.cuda.current_device())
    if is_tensor(arr):
        return arr.to(dev), lambda out: out.cpu()
    host = arr if arr.flags.writeable else arr.copy()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        src = t.from_numpy(host)
    return src.to(dev), lambda out: out.cpu().numpy()


def stream_ptr(tensor):
    t = torch()
    return ctypes.c_void_p(t.cuda.current_stream(tensor.device).cuda_stream)


def empty_like_dtype(tensor, shape):
    return torch().empty(tuple(int(s) for s in shape), dtype=tensor.dtype, device=tensor.device)


def workspace(tensor, nbytes):
    if nbytes == 0:
        return None
    return torch().empty(int(nbytes), dtype=torch().uint8, device=tensor.device)


def ptr(tensor):
    return ctypes.c_void_p(tensor.data_ptr()) if tensor is not None else ctypes.c_void_p(0)
