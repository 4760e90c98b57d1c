r"""MI355X-native approximate discrete Radon transform (drop-in for ``adrt``).

The transform core of karlotness/adrt -- ``adrt``, ``bdrt``, ``iadrt``,
``iadrt_fmg`` and the ``core`` step routines -- re-implemented as
hand-written HIP kernels for gfx950 behind a C ABI (``include/adrt_hip.h``).
Functions take NumPy arrays (host round trip) or torch tensors (device
resident) with the reference's shapes, dtypes, batch rule and errors.
There is no CPU fallback.
"""

import itertools

import numpy as np

from ._wrappers import adrt, iadrt, bdrt
from . import utils, core
from . import _device as dev
from . import _kernels

__all__ = ["adrt", "iadrt", "bdrt", "iadrt_fmg", "utils", "core"]
__version__ = "0.1.0"


def _residual_norm(x, d_a):
    """||adrt(x) - a||_2: difference in the data dtype, norm accumulated in
    float64 on the device (the batched dot kernel, whose per-image sums do
    not depend on the batch, so core.iadrt_fmg_batched decides each image
    exactly as this function does)."""
    diff = (_kernels.adrt(x) - d_a).reshape(1, -1)
    return float(_kernels.batched_dot(diff, diff)[0].sqrt())


def iadrt_fmg(a, /, *, max_iters=None):
    r"""Approximate inverse by the full multigrid method (__init__.py:69-148).

    Refines FMG estimates until the residual norm ``||adrt(x) - a||`` fails
    to decrease (or ``max_iters`` estimates were made) and returns the last
    estimate whose residual decreased.  Unbatched input only.  The iterates
    are bit-identical to the reference's; the residual norms are computed in
    float64 on the GPU, so an iterate pair whose float32 norms tie may be
    resolved differently than the reference's float32 BLAS norm would.
    """
    if a.ndim > 3:
        raise ValueError(
            f"batch dimension not supported for iadrt_fmg, got {a.ndim} dimensions"
        )
    if max_iters is not None and max_iters < 1:
        raise ValueError(f"must allow at least one iteration, but specified {max_iters}")

    state = {}

    def estimates():
        arr, _ = core._fmg_validate(a)
        d_a, restore = dev.to_device(arr)
        state["restore"] = restore
        for x in core._fmg_iter_device(d_a):
            yield x, _residual_norm(x, d_a)

    inv1 = None
    for (inv1, res1), (_, res2) in itertools.pairwise(
        itertools.chain(itertools.islice(estimates(), max_iters), [(None, np.inf)])
    ):
        if not res2 < res1:
            break
    return state["restore"](inv1.clone())
