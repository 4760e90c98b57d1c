"""Public wrappers over the native module (counterpart of src/adrt/_wrappers.py).

``_normalize_array`` follows _wrappers.py:96-112: anything that is not an
array is rejected with the reference's message; NumPy arrays are converted to
native byte order and C order (copied if unaligned), so public functions
accept Fortran-order, strided, byte-swapped and unaligned input.  torch
tensors are made contiguous.  Steps pass through ``operator.index``
(_wrappers.py:202, 325).
"""

import operator

import numpy as np

from . import _adrt_cdefs
from . import _device as dev

__all__ = []


def _set_module(module):
    def decorate(func):
        func.__module__ = module
        return func

    return decorate


def _type_name(obj, /):
    """The name of ``obj``'s type as the reference's messages print it
    (_wrappers.py:82-93): the qualified name alone for built-in types, else
    prefixed by its module."""
    cls = type(obj)
    if cls.__module__ == "builtins":
        return cls.__qualname__
    return ".".join((cls.__module__, cls.__qualname__))


def _normalize_array(a, /):
    """What the native layer takes: a contiguous device tensor, or a base-class
    NumPy array in C order, native byte order and aligned (copied only when
    one of those fails); anything else raises the reference's TypeError."""
    if dev.is_tensor(a):
        return a.contiguous()
    if not isinstance(a, np.ndarray):
        raise TypeError(f"array must be numpy.ndarray, but got {_type_name(a)}")
    return np.require(a, a.dtype.newbyteorder("="), ["C", "A", "E"])


@_set_module("adrt_amd")
def adrt(a, /):
    """Approximate discrete Radon transform (drop-in for ``adrt.adrt``).

    ``(..., N, N)`` float32/float64 with at most one batch dimension ->
    ``(..., 4, 2N-1, N)``.  Bit-identical to the reference.
    """
    return _adrt_cdefs.adrt(_normalize_array(a))


@_set_module("adrt_amd.utils")
def interp_to_cart(a, /):
    """Nearest-neighbour interpolation of an ADRT output onto a regular
    Cartesian (theta, t) grid (drop-in for ``adrt.utils.interp_to_cart``,
    _wrappers.py:328-360): ``(..., 4, 2N-1, N)`` float32/float64 with at
    most one batch dimension -> ``(..., N, 4N)``, bit-identical to the
    reference."""
    return _adrt_cdefs.interp_to_cart(_normalize_array(a))


@_set_module("adrt_amd.core")
def adrt_step(a, /, step):
    """One forward ADRT step (drop-in for ``adrt.core.adrt_step``)."""
    return _adrt_cdefs.adrt_step(_normalize_array(a), operator.index(step))


@_set_module("adrt_amd")
def iadrt(a, /):
    """Exact inverse of the ADRT (drop-in for ``adrt.iadrt``)."""
    return _adrt_cdefs.iadrt(_normalize_array(a))


@_set_module("adrt_amd")
def bdrt(a, /):
    """Back-projection (drop-in for ``adrt.bdrt``).

    ``truncate(bdrt(a)).mean(axis=-3)`` is the docs' transpose (A^T / 4);
    ``truncate(bdrt(a)).sum(axis=-3)`` is the exact adjoint A^T.
    """
    return _adrt_cdefs.bdrt(_normalize_array(a))


@_set_module("adrt_amd.core")
def bdrt_step(a, /, step):
    """One bdrt step (drop-in for ``adrt.core.bdrt_step``)."""
    return _adrt_cdefs.bdrt_step(_normalize_array(a), operator.index(step))


@_set_module("adrt_amd.core")
def threading_enabled():
    """True: the core routines split their work across many GPU threads."""
    return _adrt_cdefs.OPENMP_ENABLED


def _press_fmg_restriction(a, /):
    return _adrt_cdefs.press_fmg_restriction(_normalize_array(a))


def _press_fmg_prolongation(a, /):
    return _adrt_cdefs.press_fmg_prolongation(_normalize_array(a))


def _press_fmg_highpass(a, /):
    return _adrt_cdefs.press_fmg_highpass(_normalize_array(a))
