"""The reference's own native core, compiled -- TEST INFRASTRUCTURE ONLY.

Binds ``oracle/_ref/libadrt_ref.so``, which ``oracle/Makefile`` compiles from
the reference checkout (src/adrt/adrt_cdefs_*.hpp and adrt_cdefs_common.cpp,
where they lie) with ``-O3 -fopenmp -DNDEBUG`` plus the C-ABI shim
``oracle/ref_shim.cpp``.  Used to pin the oracle restatement bit for bit and
as the CPU baseline timed by bench.py.  ``available()`` is False when the
library was never built (no reference checkout at build time).
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libadrt_ref.so")
_lib = None


def available():
    return os.path.exists(LIB_PATH)


def _load():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"reference library not built at {LIB_PATH}")
        lib = ctypes.CDLL(LIB_PATH)
        sz = ctypes.c_size_t
        vp = ctypes.c_void_p
        for suf in ("f32", "f64"):
            for name in ("adrt", "bdrt", "iadrt", "fmg_restriction", "interp"):
                fn = getattr(lib, f"ref_{name}_{suf}")
                fn.argtypes = [vp, sz, sz, vp]
                fn.restype = ctypes.c_int
            for name in ("adrt_step", "bdrt_step"):
                fn = getattr(lib, f"ref_{name}_{suf}")
                fn.argtypes = [vp, sz, sz, ctypes.c_int, vp]
                fn.restype = ctypes.c_int
            for name in ("fmg_prolongation", "fmg_highpass"):
                fn = getattr(lib, f"ref_{name}_{suf}")
                fn.argtypes = [vp, sz, sz, sz, vp]
                fn.restype = ctypes.c_int
        lib.ref_openmp_enabled.restype = ctypes.c_int
        _lib = lib
    return _lib


def _suf(a):
    return {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64"}[a.dtype]


def _run(name, a, out_shape, *args):
    a = np.ascontiguousarray(a)
    out = np.empty(out_shape, dtype=a.dtype)
    rc = getattr(_load(), f"ref_{name}_{_suf(a)}")(a.ctypes.data, *args, out.ctypes.data)
    if rc != 0:
        raise MemoryError(name)
    return out


def _batch(a, img_dims):
    return a.shape[0] if a.ndim == img_dims + 1 else 1


def adrt(a):
    n = a.shape[-1]
    return _run("adrt", a, (*a.shape[:-2], 4, 2 * n - 1, n), _batch(a, 2), n)


def bdrt(a):
    return _run("bdrt", a, a.shape, _batch(a, 3), a.shape[-1])


def iadrt(a):
    return _run("iadrt", a, a.shape, _batch(a, 3), a.shape[-1])


def adrt_step(a, step):
    return _run("adrt_step", a, a.shape, _batch(a, 3), a.shape[-1], int(step))


def bdrt_step(a, step):
    return _run("bdrt_step", a, a.shape, _batch(a, 3), a.shape[-1], int(step))


def fmg_restriction(a):
    n = a.shape[-1]
    return _run("fmg_restriction", a, (*a.shape[:-2], n - 1, n // 2), _batch(a, 3), n)


def fmg_prolongation(a):
    h, w = a.shape[-2:]
    return _run("fmg_prolongation", a, (*a.shape[:-2], 2 * h, 2 * w), _batch(a, 2), h, w)


def fmg_highpass(a):
    h, w = a.shape[-2:]
    return _run("fmg_highpass", a, a.shape, _batch(a, 2), h, w)


def openmp_enabled():
    return bool(_load().ref_openmp_enabled())


def interp_to_cart(a):
    """The reference's interp_adrtcart (float index math): (B?, 4, 2n-1, n)
    -> (B?, n, 4n)."""
    n = a.shape[-1]
    return _run("interp", a, (*a.shape[:-3], n, 4 * n), _batch(a, 3), n)
