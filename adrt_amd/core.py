"""Lower-level routines (counterpart of src/adrt/core.py).

Step-wise transforms, their iterators, and the Press full-multigrid (FMG)
inverse steps.  All compute runs on the GPU; NumPy inputs make one host
round trip per public call and the FMG orchestration stays on the device in
between.
"""

import operator

import numpy as np

from . import _device as dev
from . import _kernels
from ._wrappers import (
    _type_name,
    adrt_step,
    bdrt_step,
    threading_enabled,
    _normalize_array,
)

__all__ = [
    "num_iters",
    "adrt_step",
    "adrt_init",
    "adrt_iter",
    "bdrt_step",
    "bdrt_iter",
    "threading_enabled",
    "iadrt_fmg_step",
    "iadrt_fmg_iter",
    "iadrt_fmg_batched",
]


def num_iters(n, /):
    r"""Number of ADRT iterations for an image of size n (core.py:94-120).

    ``ceil(log2(n))`` with ``num_iters(0) == 0``; negative values raise
    ``ValueError``, non-integers ``TypeError``.
    """
    n = operator.index(n)
    if n < 0:
        raise ValueError(f"non-negative value required for iteration count, but got {n}")
    return n.bit_length() - (bin(n).count("1") == 1)


def adrt_init(a, /):
    r"""Initialize an array for :func:`adrt_step` (core.py:123-176).

    Any dtype with 1, 2, 4 or 8-byte elements.  ``(..., N, N)`` ->
    ``(..., 4, 2N-1, N)`` with the quadrant flips in the top N offsets.
    """
    if not (isinstance(a, np.ndarray) or dev.is_tensor(a)):
        raise TypeError(f"array must be numpy.ndarray, but got {_type_name(a)}")
    if a.ndim > 3 or a.ndim < 2:
        raise ValueError(f"array must have between 2 and 3 dimensions, but had {a.ndim}")
    if a.shape[-1] != a.shape[-2] or bin(operator.index(a.shape[-1])).count("1") != 1:
        raise ValueError("array must be square with a power of two shape")
    if not all(a.shape):
        raise ValueError(
            "all array dimensions must be nonzero, "
            f"but found zero in dimension {tuple(a.shape).index(0)}"
        )
    if dev.element_size(a) not in (1, 2, 4, 8):
        raise TypeError(f"unsupported array dtype {dev.dtype_name(a)}")
    t = _normalize_array(a)
    d_in, restore = dev.to_device(t)
    return restore(_kernels.adrt_init(d_in))


def _freeze(x, copy):
    """Yield-time view/copy semantics of core.py:217-223 for NumPy results;
    tensors are cloned when copy=True and yielded as-is otherwise."""
    if dev.is_tensor(x):
        return x.clone() if copy else x
    x.setflags(write=False)
    return x.copy() if copy else x.view()


def adrt_iter(a, /, *, copy=True):
    r"""Yield the initialized array, then the output of every ADRT step
    (core.py:179-223); num_iters(N) + 1 items."""
    a = adrt_init(a)
    yield _freeze(a, copy)
    for i in range(num_iters(a.shape[-1])):
        a = adrt_step(a, i)
        yield _freeze(a, copy)


def bdrt_iter(a, /, *, copy=True):
    r"""Yield the output of every bdrt step (core.py:226-262); num_iters(N) items."""
    for i in range(num_iters(a.shape[-1])):
        a = bdrt_step(a, i)
        yield _freeze(a, copy)


def _fmg_step_device(t):
    """Press FMG inverse estimate on a device tensor (core.py:265-331).

    At level n the correction is
    ``highpass(mean_q(truncate(bdrt(adrt(ret) - a_n)) / (n - 1)))``; bdrt is
    computed for the rows truncate keeps only, and the fused truncate_combine
    kernel divides per element, sums ((q0+q1)+q2)+q3 and divides by 4,
    matching NumPy's mean over the quadrant axis.
    """
    stack = []
    for _ in range(num_iters(t.shape[-1])):
        stack.append(t)
        t = _kernels.fmg_restriction(t)
    ret = t[..., 0, :, :].contiguous()
    n = 1
    while stack:
        n *= 2
        ret = _kernels.fmg_prolongation(ret)
        resid = _kernels.adrt(ret) - stack.pop()
        ret -= _kernels.fmg_highpass(
            _kernels.adjoint(resid, divisor=n - 1, mode=_kernels.MEAN)
        )
    return ret


def _fmg_validate(a):
    # The first native call of the reference is the restriction (or, for
    # N == 1, nothing): it raises the same errors for bad input.
    from . import _adrt_cdefs

    arr = _adrt_cdefs._extract_array(_normalize_array(a))
    shape = _adrt_cdefs._array_shape(arr, 3, 4)
    n = shape[3]
    from . import _lib

    lib = _lib.load()
    if not lib.adrt_hip_valid_output_shape(_lib.shape4(shape)):
        raise ValueError("array must have a valid ADRT output shape")
    _adrt_cdefs._dtype_code(arr)
    return arr, n


def iadrt_fmg_step(a, /):
    r"""Estimated inverse by the full multigrid method (core.py:265-331).

    Accepts an ADRT output with an optional batch dimension; returns the
    ``(..., N, N)`` estimate.
    """
    arr, _ = _fmg_validate(a)
    d_in, restore = dev.to_device(arr)
    return restore(_fmg_step_device(d_in))


def _fmg_iter_device(d_a):
    """Device-resident refinement loop.  Calls the module-level
    ``iadrt_fmg_step`` (so wrappers installed on it see every step) with
    device tensors, which it returns unchanged in kind."""
    inv = iadrt_fmg_step(d_a)
    yield inv
    while True:
        inv = inv + iadrt_fmg_step(d_a - _kernels.adrt(inv))
        yield inv


def iadrt_fmg_iter(a, /, *, copy=True):
    r"""Infinite iterator of refined FMG inverses (core.py:334-381):
    ``inv_{k+1} = inv_k + iadrt_fmg_step(a - adrt(inv_k))``.

    The iterates live on the GPU; NumPy callers receive a host copy of each
    (read-only view semantics for ``copy=False`` as in the reference).
    """
    arr, _ = _fmg_validate(a)
    d_a, restore = dev.to_device(arr)
    for inv in _fmg_iter_device(d_a):
        yield _freeze(restore(inv), copy)


def iadrt_fmg_batched(a, /, *, max_iters=None):
    r"""Batched :func:`adrt_amd.iadrt_fmg`: ``(B, 4, 2N-1, N)`` ->
    ``(B, N, N)``.

    An extension: the reference's ``iadrt_fmg`` takes one image
    (``__init__.py:122-125``).  Every image follows that function's rule on
    its own: its FMG estimates are refined (``iadrt_fmg_iter``) until its
    residual norm ``||adrt(x) - a||`` fails to decrease, or ``max_iters``
    estimates were made, and the last estimate whose residual decreased is
    returned.  The refinement runs on the whole batch on the GPU until every
    image has stopped; image ``b`` of the result equals ``iadrt_fmg(a[b])``
    bit for bit (the estimates and the float64 residual norms of an image do
    not depend on the rest of the batch).
    """
    import itertools

    if a.ndim != 4:
        raise ValueError(f"array must have 4 dimensions (a batch), but had {a.ndim}")
    if max_iters is not None and max_iters < 1:
        raise ValueError(f"must allow at least one iteration, but specified {max_iters}")
    arr, _ = _fmg_validate(a)
    d_a, restore = dev.to_device(arr)
    torch = dev.torch()
    B = d_a.shape[0]
    best = best_res = None
    done = torch.zeros(B, dtype=torch.bool, device=d_a.device)
    for x in itertools.islice(_fmg_iter_device(d_a), max_iters):
        diff = (_kernels.adrt(x) - d_a).reshape(B, -1)
        res = _kernels.batched_dot(diff, diff).sqrt()
        if best is None:
            best, best_res = x.clone(), res
            continue
        better = ~done & (res < best_res)
        done |= ~better
        best[better] = x[better]
        best_res = torch.where(better, res, best_res)
        if bool(done.all()):
            break
    return restore(best)
