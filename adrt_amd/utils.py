"""Utilities (counterpart of src/adrt/utils.py).

``truncate`` (utils.py:191-242) runs as a gfx950 copy kernel for every
element size of 1, 2, 4 or 8 bytes, on NumPy arrays (host round trip) and
torch tensors, with any number of leading dimensions and any base size N.
``interp_to_cart`` (_wrappers.py:328-360) is the nearest-neighbour
resampling kernel; ``stitch_adrt`` / ``unstitch_adrt`` (utils.py:65-188) are
gather kernels with the reference's index rules, for the same element sizes.
``coord_adrt`` / ``coord_cart_to_adrt`` (utils.py:245-435) are host-side
coordinate maps (NumPy, float64), as in the reference.
"""

import operator
import typing

import numpy as np

from . import _device as dev
from . import _kernels
from ._wrappers import interp_to_cart

__all__ = ["stitch_adrt", "unstitch_adrt", "truncate", "coord_adrt", "coord_cart_to_adrt",
           "interp_to_cart"]


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


def _shape_str(a):
    return str(tuple(int(d) for d in a.shape)) if dev.is_tensor(a) else str(a.shape)


def _gather(a, out_shape, fn_name, outer, n, flag):
    """Run a layout gather kernel on a NumPy array (host round trip) or a
    device tensor."""
    from . import _lib

    t = a if dev.is_tensor(a) else np.asarray(a)
    _check_copyable(t)
    if not dev.is_tensor(t):
        t = np.ascontiguousarray(t)
        if not t.dtype.isnative:
            t = t.astype(t.dtype.newbyteorder("="))
    d_in, restore = dev.to_device(t)
    d_in = d_in.contiguous()
    out = dev.empty_like_dtype(d_in, out_shape)
    if out.numel() == 0:
        return restore(out)
    lib = _lib.load()
    _lib.check(getattr(lib, fn_name)(dev.ptr(d_in), dev.ptr(out), d_in.element_size(), outer, n,
                                     int(flag), dev.stream_ptr(d_in)), fn_name)
    return restore(out)


def stitch_adrt(a, /, *, remove_repeated=False):
    r"""Reshape and align ADRT quadrants into one contiguous image (drop-in
    for ``adrt.utils.stitch_adrt``, utils.py:65-130).

    ``(..., 4, 2N-1, N)`` (any N, any leading dimensions) ->
    ``(..., 3N-2, 4N)``, or ``(..., 3N-2, 4N-4)`` with `remove_repeated`
    (the last column of each quadrant dropped).  Quadrants stand side by
    side, the odd ones flipped in both axes; quadrants 0 and 1 fill the top
    2N-1 rows, 2 and 3 the bottom 2N-1; the rest is zero.
    """
    n = a.shape[-1] if len(a.shape) >= 1 else None
    if len(a.shape) < 3 or tuple(a.shape[-3:]) != (4, 2 * n - 1, n):
        raise ValueError(f"unsuitable shape for ADRT output processing {_shape_str(a)}")
    vc = n - (1 if remove_repeated else 0)
    outer = int(np.prod(a.shape[:-3], dtype=np.int64)) if len(a.shape) > 3 else 1
    out_shape = (*a.shape[:-3], 3 * n - 2, 4 * vc)
    if vc == 0 or outer == 0:
        return _gather(a, out_shape, "adrt_hip_stitch", 0, n, False)
    return _gather(a, out_shape, "adrt_hip_stitch", outer, n, bool(remove_repeated))


def unstitch_adrt(a, /):
    r"""Slice a stitched ADRT output back into its quadrants (drop-in for
    ``adrt.utils.unstitch_adrt``, utils.py:133-188): ``(..., 3N-2, 4N)`` or
    ``(..., 3N-2, 4N-4)`` -> ``(..., 4, 2N-1, N)``.  Columns removed by
    ``stitch_adrt(..., remove_repeated=True)`` are restored from the
    neighbouring quadrant, as the reference does.
    """
    if len(a.shape) < 2:
        raise ValueError(f"unsuitable shape for ADRT unstitching {_shape_str(a)}")
    n = (a.shape[-2] + 2) // 3
    if a.shape[-2] != 3 * n - 2 or (a.shape[-1] != 4 * n and a.shape[-1] != 4 * n - 4):
        raise ValueError(f"unsuitable shape for ADRT unstitching {_shape_str(a)}")
    removed = a.shape[-1] == 4 * n - 4
    outer = int(np.prod(a.shape[:-2], dtype=np.int64)) if len(a.shape) > 2 else 1
    out_shape = (*a.shape[:-2], 4, 2 * n - 1, n)
    return _gather(a, out_shape, "adrt_hip_unstitch", outer, n, removed)


class ADRTCoord(typing.NamedTuple):
    offset: np.ndarray
    angle: np.ndarray


class ADRTIndex(typing.NamedTuple):
    quadrant: np.ndarray
    height: np.ndarray
    slope: np.ndarray
    factor: np.ndarray


def _domain_size(n):
    n = operator.index(n)
    if n < 2:
        raise ValueError(f"invalid Radon domain size {n}, must be at least 2")
    if n & (n - 1):
        raise ValueError(f"invalid Radon domain size {n}, must be a power of two")
    return n


def coord_adrt(n, /):
    r"""Radon-domain coordinates of every ADRT output entry (drop-in for
    ``adrt.utils.coord_adrt``, utils.py:245-331).

    Returns ``(offset, angle)``: ``offset`` (4, 2n-1, n) float64 and
    ``angle`` (4, 1, n) float64 in [-pi/2, pi/2].  Digital line k of a
    quadrant has slope k/(n-1), i.e. base angle atan(k/(n-1)); row h sits at
    height 1 - (h + 1/2)/n on the quadrant's axis, and its signed distance to
    the centre is ((height + (2n-1)/(2n) * slope) / (1 + slope) - 1/2) *
    (cos + sin) of the base angle.  Quadrants 1 and 3 mirror the offsets;
    the angles of quadrants 0..3 are base - pi/2, -base, base, -(base - pi/2).

    The slopes and heights are built as the reference builds them, with
    NumPy's ``linspace`` (slope k = k * fl(1/(n-1)), last entry exactly 1;
    height h = 1 + h * step + step/2 with step = ((1-n)/n - 1)/(2n-1)), and
    every later operation is the reference's in its order, so the values are
    the reference's bit for bit (k/(n-1) rounds differently for some k).
    """
    n = _domain_size(n)
    height, step = np.linspace(1, (1 - n) / n, num=2 * n - 1, endpoint=False, retstep=True,
                               dtype=np.float64)
    height += step / 2
    slope = np.linspace(0, 1, num=n, endpoint=True, dtype=np.float64)
    base = np.arctan(slope)
    line_pos = np.add.outer(height, ((2 * n - 1) / (2 * n)) * slope) / (1 + slope)
    dist = (line_pos - 0.5) * (np.cos(base) + np.sin(base))
    offset = np.stack([dist, -dist, dist, -dist])
    below = base - np.pi / 2
    angle = np.stack([below, -base, base, -below])[:, None, :]
    return ADRTCoord(offset, angle)


def coord_cart_to_adrt(theta, t, n):
    r"""Nearest ADRT entries for continuous Radon points (theta, t) (drop-in
    for ``adrt.utils.coord_cart_to_adrt``, utils.py:340-435).

    Returns ``(quadrant, height, slope, factor)`` arrays shaped like the
    inputs (uint8, int64, uint64, float64): theta is first brought into
    [-pi/2, pi/2] (period pi); the quadrant is floor(theta/(pi/4)) + 2
    clipped to 0..3 (on a quadrant boundary the lower quadrant); with th0 the
    distance of theta to the nearest multiple of pi/2, the slope index is
    tan(th0)*(n-1) rounded half to even, the factor sqrt(1 + (slope/(n-1))^2),
    and the height index floor((round_half_even(2h) - 1)/2) for
    h = n * ((1 + tan th0)/2 + s * t / cos th0), s = +1 in quadrants 1, 3 and
    -1 in 0, 2.  Computed in the inputs' precision, as NumPy promotes them.
    """
    n = _domain_size(n)
    theta = np.asarray(theta)
    t = np.asarray(t)
    if theta.shape != t.shape:
        raise ValueError(f"mismatched shapes for theta and t {theta.shape} vs. {t.shape}")
    inside = np.abs(theta) <= np.pi / 2
    theta = np.where(inside, theta, np.remainder(theta + np.pi / 2, np.pi) - np.pi / 2)
    quadrant = np.floor(np.clip(theta / (np.pi / 4), -2, 1)).astype(np.int8) + 2
    th0 = np.pi / 4 - np.abs(np.abs(theta) - np.pi / 4)
    tan0 = np.tan(th0)
    slope = np.around(tan0 * (n - 1)).astype(np.uint64)
    factor = np.sqrt(1 + (slope / (n - 1)) ** 2)
    side = 2 * (quadrant % 2) - 1
    h = (0.5 * (1 + tan0) + (side * t) / np.cos(th0)) * n
    height = (np.round(2 * h).astype(np.int64) - 1) // 2
    return ADRTIndex(quadrant.astype(np.uint8), height, slope, factor)
