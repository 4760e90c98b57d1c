"""Utilities (counterpart of src/adrt/utils.py).

``truncate`` (utils.py:191-242) runs as a gfx950 copy kernel for every
element size of 1, 2, 4 or 8 bytes, on NumPy arrays (host round trip) and
torch tensors, with any number of leading dimensions and any base size N.
``interp_to_cart`` (_wrappers.py:328-360) is the nearest-neighbour
resampling kernel.
"""

import numpy as np

from . import _device as dev
from . import _kernels
from ._wrappers import interp_to_cart

__all__ = ["truncate", "interp_to_cart"]


def _check_copyable(a):
    if dev.element_size(a) not in (1, 2, 4, 8):
        raise TypeError(f"unsupported array dtype {dev.dtype_name(a)} (element size must be 1, 2, 4 or 8 bytes)")


def truncate(a, /):
    r"""Truncate and rotate a rectangular ADRT output into a square.

    Same semantics as ``adrt.utils.truncate``: input ``(..., 4, 2N-1, N)``,
    output ``(..., 4, N, N)`` with each quadrant un-flipped into the
    orientation of the original image (the inverse of
    :func:`adrt_amd.core.adrt_init`).
    """
    if dev.is_tensor(a):
        t = a
    else:
        t = np.asarray(a)
    n = t.shape[-1] if t.ndim >= 1 else None
    if t.ndim < 3 or tuple(t.shape[-3:]) != (4, 2 * n - 1, n):
        raise ValueError(f"unsuitable shape for ADRT output processing {tuple(a.shape)}")
    _check_copyable(t)
    if not dev.is_tensor(t):
        t = np.ascontiguousarray(t)
        if not t.dtype.isnative:
            t = t.astype(t.dtype.newbyteorder("="))
    else:
        t = t.contiguous()
    d_in, restore = dev.to_device(t)
    return restore(_kernels.truncate(d_in))
