"""Drop-in counterpart of the reference's native module ``adrt._adrt_cdefs``.

Same functions, same positional arguments, same check order and the same
exception types and messages as ``src/adrt/adrt_cdefs_py.cpp`` (method table
:852-863), with the compute routed to the gfx950 HIP library through its C
ABI.  Accepted arrays: C-contiguous, aligned, native-byte-order NumPy arrays
(host round trip) and contiguous torch tensors (CUDA/HIP tensors stay on the
device).  Every call returns a new array of the input's type and dtype.

Check order of every function (adrt_cdefs_py.cpp:275-341 and siblings):
  arity -> array type -> layout -> ndim -> zero dims -> shape -> size
  overflow -> step (int, then range) -> dtype -> compute.
"""

import operator

import numpy as np

from . import _device as dev
from . import _kernels
from . import _lib

__all__ = [
    "adrt",
    "adrt_step",
    "iadrt",
    "bdrt",
    "bdrt_step",
    "press_fmg_restriction",
    "press_fmg_prolongation",
    "press_fmg_highpass",
    "OPENMP_ENABLED",
]

# The reference reports whether its core routines split work across threads
# (adrt_cdefs_common.hpp:103-108).  Here every routine runs on thousands of
# GPU threads, so the answer is yes.
OPENMP_ENABLED: bool = True

_LAYOUT_MSG = "array must be C-order, contiguous, aligned, and native byte order"
_INT_MAX = 2**31 - 1
_INT_MIN = -(2**31)
_SIZE_MAX = 2**64 - 1


def _extract_array(arg):
    """adrt_cdefs_py.cpp:77-90."""
    if isinstance(arg, np.ndarray):
        if not (arg.flags.c_contiguous and arg.flags.aligned and arg.dtype.isnative):
            raise ValueError(_LAYOUT_MSG)
        return arg
    if dev.is_tensor(arg):
        if not arg.is_contiguous():
            raise ValueError(_LAYOUT_MSG)
        return arg
    raise TypeError("array must be a NumPy array or compatible subclass")


def _array_shape(arr, min_dim, max_dim):
    """adrt_cdefs_py.cpp:121-154: ndim range, then nonzero dims, padded with
    leading ones to max_dim."""
    nd = arr.ndim
    if nd < min_dim or nd > max_dim:
        raise ValueError(
            f"array must have between {min_dim} and {max_dim} dimensions, but had {nd}"
        )
    shape = tuple(int(s) for s in arr.shape)
    for i, s in enumerate(shape):
        if s <= 0:
            raise ValueError(
                f"all array dimensions must be nonzero, but found zero in dimension {i}"
            )
    return (1,) * (max_dim - nd) + shape


def _check_size(shape):
    """adrt_cdefs_py.cpp:190-197: element count must fit size_t."""
    prod = 1
    for s in shape:
        prod *= s
    if prod > _SIZE_MAX:
        raise ValueError("array is too big; unable to allocate temporary space")


def _extract_int(obj):
    """adrt_cdefs_py.cpp:100-119: PyLong_AsLong semantics narrowed to C int."""
    val = operator.index(obj)
    if not (_INT_MIN <= val <= _INT_MAX):
        raise OverflowError("Python int too large to convert to C int")
    return val


def _dtype_code(arr):
    code = dev.dtype_code(arr)
    if code is None:
        raise TypeError(f"unsupported array dtype {dev.dtype_name(arr)}")
    return code


def _unpack(args, n, name):
    if len(args) != n:
        raise TypeError(f"{name} expected {n} arguments, got {len(args)}")
    return args


def _compute(arr, fn, *args):
    """Run a device kernel on arr (host round trip for NumPy / CPU tensors)."""
    d_in, restore = dev.to_device(arr)
    return restore(fn(d_in, *args))


def adrt(a, /):
    """adrt_cdefs_py.cpp:275-341 -> adrt_hip_adrt."""
    arr = _extract_array(a)
    shape = _array_shape(arr, 2, 3)
    lib = _lib.load()
    if not lib.adrt_hip_valid_adrt_shape(_lib.shape3(shape)):
        raise ValueError("array must be square with a power of two shape")
    b, _, n = shape
    out_virtual = (b, 4, 2 * n - 1, n)
    _check_size(out_virtual)
    _dtype_code(arr)
    return _compute(arr, _kernels.adrt)


def _output_shape_op(a, min_dim=3, max_dim=4):
    arr = _extract_array(a)
    shape = _array_shape(arr, min_dim, max_dim)
    return arr, shape


def iadrt(a, /):
    """adrt_cdefs_py.cpp:414-480 -> adrt_hip_iadrt."""
    arr, shape = _output_shape_op(a)
    lib = _lib.load()
    if not lib.adrt_hip_valid_output_shape(_lib.shape4(shape)):
        raise ValueError("array must have a valid ADRT output shape")
    _check_size(shape)
    _dtype_code(arr)
    return _compute(arr, _kernels.iadrt)


def bdrt(a, /):
    """adrt_cdefs_py.cpp:482-548 -> adrt_hip_bdrt."""
    arr, shape = _output_shape_op(a)
    lib = _lib.load()
    if not lib.adrt_hip_valid_output_shape(_lib.shape4(shape)):
        raise ValueError("array must have a valid ADRT output shape")
    _check_size(shape)
    _dtype_code(arr)
    return _compute(arr, _kernels.bdrt)


def _step(name, shape_msg, kernel, args):
    a, step = _unpack(args, 2, name)
    arr, shape = _output_shape_op(a)
    lib = _lib.load()
    if not lib.adrt_hip_valid_output_shape(_lib.shape4(shape)):
        raise ValueError(shape_msg)
    it = _extract_int(step)
    if not lib.adrt_hip_valid_step(_lib.shape4(shape), it):
        raise ValueError(
            f"step {it} is out of range for array's shape, use adrt.core.num_iters"
        )
    _dtype_code(arr)
    return _compute(arr, kernel, it)


def adrt_step(*args):
    """adrt_cdefs_py.cpp:343-412 -> adrt_hip_adrt_step."""
    return _step(
        "adrt_step",
        "array must have valid shape for ADRT, use adrt.core.adrt_init",
        _kernels.adrt_step,
        args,
    )


def bdrt_step(*args):
    """adrt_cdefs_py.cpp:550-619 -> adrt_hip_bdrt_step."""
    return _step(
        "bdrt_step", "array must have a valid ADRT output shape", _kernels.bdrt_step, args
    )


def press_fmg_restriction(a, /):
    """adrt_cdefs_py.cpp:684-738 -> adrt_hip_fmg_restriction."""
    arr, shape = _output_shape_op(a)
    lib = _lib.load()
    if not lib.adrt_hip_valid_restriction_shape(_lib.shape4(shape)):
        raise ValueError("array must have a valid ADRT output shape")
    _dtype_code(arr)
    return _compute(arr, _kernels.fmg_restriction)


def press_fmg_prolongation(a, /):
    """adrt_cdefs_py.cpp:740-794 -> adrt_hip_fmg_prolongation."""
    arr = _extract_array(a)
    shape = _array_shape(arr, 2, 3)
    lib = _lib.load()
    if not lib.adrt_hip_valid_prolongation_shape(_lib.shape3(shape)):
        raise ValueError("array is too large for prolongation operator")
    _dtype_code(arr)
    return _compute(arr, _kernels.fmg_prolongation)


def press_fmg_highpass(a, /):
    """adrt_cdefs_py.cpp:796-850 -> adrt_hip_fmg_highpass."""
    arr = _extract_array(a)
    shape = _array_shape(arr, 2, 3)
    lib = _lib.load()
    if not lib.adrt_hip_valid_highpass_shape(_lib.shape3(shape)):
        raise ValueError("array is too small to high-pass filter")
    _dtype_code(arr)
    return _compute(arr, _kernels.fmg_highpass)


def interp_to_cart(a, /):
    """adrt_cdefs_py.cpp:621-682 -> adrt_hip_interp_to_cart: the shape rule of
    bdrt plus n > 1 (adrt_cdefs_common.cpp:276-279), then the float index
    limit (4n exact in float), then the dtype."""
    arr, shape = _output_shape_op(a)
    lib = _lib.load()
    if not lib.adrt_hip_valid_output_shape(_lib.shape4(shape)) or shape[3] <= 1:
        raise ValueError("array must have a valid ADRT output shape")
    if shape[3] > (1 << 24) // 4:
        raise ValueError("array is too big for interpolation index calculations")
    _dtype_code(arr)
    return _compute(arr, _kernels.interp_to_cart)
