"""CPU oracle for the ADRT transform core -- TEST INFRASTRUCTURE ONLY.

NumPy front end over ``liboracle.so`` (the loop-for-loop C restatement of the
reference's native routines) plus NumPy restatements of the reference's
pure-Python routines.  Only tests, the smoke check and bench.py's CPU
baseline leg use this module.  Each function names the reference code it
follows; how each is pinned is listed in the test files under ``tests/``.

Pinning summary (what checks each function beyond its own restatement):

* adrt / adrt_step: bit-equality with the compiled reference
  (``oracle/_ref``) and the committed golden vectors; brute-force digital-line
  sums on integer images; spot checks of tests/test_adrt.py.
* bdrt / bdrt_step: bit-equality with ``oracle/_ref`` and golden vectors;
  adjointness with adrt on materialised matrices.
* iadrt: bit-equality with ``oracle/_ref`` and golden vectors; exact round
  trip on integer images.
* fmg_*: bit-equality with ``oracle/_ref``; NumPy closed forms; constant /
  checkerboard fixed points of the high-pass filter.
* adrt_init / truncate: truncate(adrt_init(x)) == x; quadrant formulas.
* iadrt_fmg_step / iadrt_fmg: a recursive Press-FMG restatement and round
  trips (tolerance 1e-3 as in the reference tests).
* normal operator / CG: symmetry and positive definiteness; parity with the
  docs recipe is unpinned (the recipe has no reference test).
"""

import ctypes
import itertools
import operator
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(
                f"oracle library missing at {_LIB_PATH}; run `make -C oracle`"
            )
        lib = ctypes.CDLL(_LIB_PATH)
        sz = ctypes.c_size_t
        vp = ctypes.c_void_p
        for suf in ("f32", "f64"):
            for name in ("adrt", "bdrt", "iadrt", "fmg_restriction"):
                fn = getattr(lib, f"oracle_{name}_{suf}")
                fn.argtypes = [vp, sz, sz, vp]
                fn.restype = ctypes.c_int
            for name in ("adrt_step", "bdrt_step"):
                fn = getattr(lib, f"oracle_{name}_{suf}")
                fn.argtypes = [vp, sz, sz, ctypes.c_int, vp]
                fn.restype = ctypes.c_int
            for name in ("fmg_prolongation", "fmg_highpass"):
                fn = getattr(lib, f"oracle_{name}_{suf}")
                fn.argtypes = [vp, sz, sz, sz, vp]
                fn.restype = ctypes.c_int
        lib.oracle_openmp_enabled.restype = ctypes.c_int
        _lib = lib
    return _lib


def _suffix(a):
    if a.dtype == np.float32:
        return "f32"
    if a.dtype == np.float64:
        return "f64"
    raise TypeError(f"unsupported array dtype {a.dtype}")


def _prep(a, min_dim, max_dim):
    a = np.ascontiguousarray(a)
    if not (min_dim <= a.ndim <= max_dim):
        raise ValueError(
            f"array must have between {min_dim} and {max_dim} dimensions, but had {a.ndim}"
        )
    return a


def _call(name, a, out_shape, *args):
    lib = _load()
    out = np.empty(out_shape, dtype=a.dtype)
    rc = getattr(lib, f"oracle_{name}_{_suffix(a)}")(
        a.ctypes.data, *args, out.ctypes.data
    )
    if rc != 0:
        raise MemoryError(f"oracle {name} failed to allocate")
    return out


def openmp_enabled():
    return bool(_load().oracle_openmp_enabled())


# --------------------------------------------------------------------------
# Native restatements (adrt_cdefs_*.hpp)
# --------------------------------------------------------------------------


def adrt(a):
    """Forward ADRT; adrt_cdefs_adrt.hpp:100-212 via liboracle."""
    a = _prep(a, 2, 3)
    n = a.shape[-1]
    assert a.shape[-2] == n and n & (n - 1) == 0
    b = a.shape[0] if a.ndim == 3 else 1
    out = _call("adrt", a, (b, 4, 2 * n - 1, n), b, n)
    return out if a.ndim == 3 else out[0]


def adrt_step(a, step):
    """One forward stage; adrt_cdefs_adrt.hpp:214-258 via liboracle."""
    a = _prep(a, 3, 4)
    n = a.shape[-1]
    b = a.shape[0] if a.ndim == 4 else 1
    return _call("adrt_step", a, a.shape, b, n, int(step))


def bdrt(a):
    """Back-projection; adrt_cdefs_bdrt.hpp:120-187 via liboracle."""
    a = _prep(a, 3, 4)
    n = a.shape[-1]
    b = a.shape[0] if a.ndim == 4 else 1
    return _call("bdrt", a, a.shape, b, n)


def bdrt_step(a, step):
    """One back-projection stage; adrt_cdefs_bdrt.hpp:189-244 via liboracle."""
    a = _prep(a, 3, 4)
    n = a.shape[-1]
    b = a.shape[0] if a.ndim == 4 else 1
    return _call("bdrt_step", a, a.shape, b, n, int(step))


def iadrt(a):
    """Exact inverse; adrt_cdefs_iadrt.hpp:109-176 via liboracle."""
    a = _prep(a, 3, 4)
    n = a.shape[-1]
    b = a.shape[0] if a.ndim == 4 else 1
    return _call("iadrt", a, a.shape, b, n)


def fmg_restriction(a):
    """Multigrid restriction; adrt_cdefs_fmg.hpp:53-73 via liboracle."""
    a = _prep(a, 3, 4)
    n = a.shape[-1]
    b = a.shape[0] if a.ndim == 4 else 1
    return _call("fmg_restriction", a, (*a.shape[:-2], n - 1, n // 2), b, n)


def fmg_prolongation(a):
    """Multigrid prolongation; adrt_cdefs_fmg.hpp:75-95 via liboracle."""
    a = _prep(a, 2, 3)
    h, w = a.shape[-2:]
    b = a.shape[0] if a.ndim == 3 else 1
    return _call("fmg_prolongation", a, (*a.shape[:-2], 2 * h, 2 * w), b, h, w)


def fmg_highpass(a):
    """Multigrid high-pass filter; adrt_cdefs_fmg.hpp:97-171 via liboracle."""
    a = _prep(a, 2, 3)
    h, w = a.shape[-2:]
    b = a.shape[0] if a.ndim == 3 else 1
    return _call("fmg_highpass", a, a.shape, b, h, w)


# --------------------------------------------------------------------------
# Restatements of the reference's pure-Python routines
# --------------------------------------------------------------------------


def num_iters(n):
    """core.py:94-120: bit_length(n) - (popcount(n) == 1)."""
    n = operator.index(n)
    if n < 0:
        raise ValueError(f"non-negative value required for iteration count, but got {n}")
    return n.bit_length() - (bin(n).count("1") == 1)


def adrt_init(a):
    """core.py:123-176: quadrant copies into the top N offsets, zeros below."""
    n = a.shape[-1]
    out = np.zeros((*a.shape[:-2], 4, 2 * n - 1, n), dtype=a.dtype)
    out[..., 0, :n, :] = np.flip(a, axis=-1).swapaxes(-1, -2)
    out[..., 1, :n, :] = np.flip(a, axis=-2)
    out[..., 2, :n, :] = a
    out[..., 3, :n, :] = np.flip(a, axis=(-1, -2)).swapaxes(-1, -2)
    return out


def truncate(a):
    """utils.py:191-242: un-flip the top n x n square of each quadrant."""
    n = a.shape[-1]
    if a.shape[-3:] != (4, 2 * n - 1, n):
        raise ValueError(f"unsuitable shape for ADRT output processing {a.shape}")
    return np.stack(
        [
            np.flip(a[..., 0, :n, :n], axis=-2).swapaxes(-1, -2),
            np.flip(a[..., 1, :n, :n], axis=-2),
            a[..., 2, :n, :n],
            np.flip(a[..., 3, :n, :n], axis=(-1, -2)).swapaxes(-1, -2),
        ],
        axis=-3,
    )


def iadrt_fmg_step(a):
    """core.py:265-331: one Press full-multigrid inverse estimate.

    The correction at level n is
    ``highpass(mean(truncate(bdrt(adrt(ret) - a_n)) / (n - 1), axis=-3))``:
    per-element divide first, then the NumPy mean over the quadrant axis.
    """
    stack = []
    for _ in range(num_iters(a.shape[-1])):
        stack.append(a)
        a = fmg_restriction(a)
    ret = a[..., 0, :, :]
    n = 1
    while stack:
        n *= 2
        ret = fmg_prolongation(ret)
        ret -= fmg_highpass(
            np.mean(truncate(bdrt(adrt(ret) - stack.pop())) / (n - 1), axis=-3)
        )
    return ret


def iadrt_fmg_iter(a, copy=True):
    """core.py:334-381: inv_{k+1} = inv_k + iadrt_fmg_step(a - adrt(inv_k))."""
    inv = iadrt_fmg_step(a)
    yield inv.copy() if copy else inv
    while True:
        inv = inv + iadrt_fmg_step(a - adrt(inv))
        yield inv.copy() if copy else inv


def iadrt_fmg(a, max_iters=None):
    """__init__.py:69-148: stop at the first non-decreasing residual norm."""
    if a.ndim > 3:
        raise ValueError(
            f"batch dimension not supported for iadrt_fmg, got {a.ndim} dimensions"
        )
    if max_iters is not None and max_iters < 1:
        raise ValueError(f"must allow at least one iteration, but specified {max_iters}")
    inv1 = a
    pairs = itertools.chain(
        (
            (x, float(np.linalg.norm(adrt(x) - a)))
            for x in itertools.islice(iadrt_fmg_iter(a, copy=False), max_iters)
        ),
        [(a, np.inf)],
    )
    for (inv1, res1), (_, res2) in itertools.pairwise(pairs):
        if not res2 < res1:
            break
    return inv1.copy()


def normal_op(x):
    """docs/examples.cginverse.md:40-67: x -> mean_q(truncate(bdrt(adrt(x))))."""
    return truncate(bdrt(adrt(x))).mean(axis=-3)


def adjoint(y):
    """A^T y = sum_q truncate(bdrt(y))_q (tests/test_bdrt.py:310-325)."""
    return truncate(bdrt(y)).sum(axis=-3)


def cg_normal(b, x0=None, rtol=1e-4, maxiter=1000):
    """Plain CG on the normal equations M x = rhs, M = normal_op, in float64.

    ``b`` is ADRT-domain data (4, 2N-1, N); the right-hand side is
    ``mean_q(truncate(bdrt(b)))`` as in the docs recipe.  Returns the iterate,
    the iteration count and the relative-residual history.
    """
    b = np.asarray(b, dtype=np.float64)
    rhs = truncate(bdrt(b)).mean(axis=-3)
    x = np.zeros_like(rhs) if x0 is None else np.array(x0, dtype=np.float64)
    r = rhs - normal_op(x)
    p = r.copy()
    rr = float(np.vdot(r, r))
    bn = float(np.linalg.norm(rhs))
    hist = [np.sqrt(rr) / bn]
    k = 0
    while hist[-1] > rtol and k < maxiter:
        ap = normal_op(p)
        alpha = rr / float(np.vdot(p, ap))
        x += alpha * p
        r -= alpha * ap
        rr_new = float(np.vdot(r, r))
        p = r + (rr_new / rr) * p
        rr = rr_new
        k += 1
        hist.append(np.sqrt(rr) / bn)
    return x, k, hist
