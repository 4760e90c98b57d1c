"""ctypes binding of the gfx950 HIP library ``libadrt_hip.so``.

This is the only module that touches the C ABI declared in
``include/adrt_hip.h``.  There is no CPU fallback: if the library is missing
or no HIP device is visible, compute calls raise ``RuntimeError``.  Shape
validation functions work without a device.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libadrt_hip.so")

F32 = 0
F64 = 1
OP_ADRT = 0
OP_BDRT = 1
OP_IADRT = 2

OK = 0
ERR_SHAPE = 1
ERR_STEP = 2
ERR_DTYPE = 3
ERR_TOO_BIG = 4
ERR_NULL = 5
ERR_WORKSPACE = 6
ERR_DEVICE = 7
ERR_ARG = 8
ERR_SCRATCH = 9
NOALLOC_SCRATCH_BYTES = 68 << 20

VARIANT_DEFAULT = 0
VARIANT_ALT_SPLIT = 1
VARIANT_STREAMING = 2
VARIANT_SINGLE = 3

_SZ = ctypes.c_size_t
_VP = ctypes.c_void_p
_INT = ctypes.c_int
_SHAPE3 = _SZ * 3
_SHAPE4 = _SZ * 4


class Batch(ctypes.Structure):
    """``adrt_hip_batch``: one entry of adrt_hip_adrt_bdrt_small."""
    _fields_ = [("images", ctypes.c_void_p), ("adrt", ctypes.c_void_p),
                ("bdrt", ctypes.c_void_p), ("batch", ctypes.c_size_t), ("n", ctypes.c_size_t)]


class Queue(ctypes.Structure):
    """``adrt_hip_queue``: a planned persistent queue (adrt_hip_queue_plan)."""
    _fields_ = [("state", ctypes.c_void_p), ("dtype", ctypes.c_int), ("entries", ctypes.c_int),
                ("segments", ctypes.c_int), ("counters", ctypes.c_int),
                ("lds_bytes", ctypes.c_int), ("waves", ctypes.c_int)]


# name -> (restype, argtypes); mirrors include/adrt_hip.h
SIGNATURES = {
    "adrt_hip_status_string": (ctypes.c_char_p, [_INT]),
    "adrt_hip_last_error": (ctypes.c_char_p, []),
    "adrt_hip_abi_version": (_INT, []),
    "adrt_hip_device_count": (_INT, []),
    "adrt_hip_num_iters": (_INT, [_SZ]),
    "adrt_hip_valid_adrt_shape": (_INT, [_SHAPE3]),
    "adrt_hip_valid_output_shape": (_INT, [_SHAPE4]),
    "adrt_hip_valid_step": (_INT, [_SHAPE4, _INT]),
    "adrt_hip_valid_restriction_shape": (_INT, [_SHAPE4]),
    "adrt_hip_valid_prolongation_shape": (_INT, [_SHAPE3]),
    "adrt_hip_valid_highpass_shape": (_INT, [_SHAPE3]),
    "adrt_hip_workspace_bytes": (_SZ, [_INT, _INT, _SZ, _SZ]),
    "adrt_hip_adrt": (_INT, [_VP, _VP, _VP, _SZ, _INT, _SZ, _SZ, _VP]),
    "adrt_hip_adrt_noalloc": (_INT, [_VP, _VP, _INT, _SZ, _SZ, _VP]),
    "adrt_hip_adrt_step": (_INT, [_VP, _VP, _INT, _SZ, _SZ, _INT, _VP]),
    "adrt_hip_bdrt": (_INT, [_VP, _VP, _VP, _SZ, _INT, _SZ, _SZ, _VP]),
    "adrt_hip_bdrt_rows": (_INT, [_VP, _VP, _VP, _SZ, _INT, _SZ, _SZ, _SZ, _VP]),
    "adrt_hip_bdrt_step": (_INT, [_VP, _VP, _INT, _SZ, _SZ, _INT, _VP]),
    "adrt_hip_adrt_bdrt_small": (_INT, [ctypes.POINTER(Batch), _INT, _INT, _INT, _VP]),
    "adrt_hip_queue_workspace_bytes": (_SZ, [ctypes.POINTER(Batch), _INT, _INT]),
    "adrt_hip_queue_state_bytes": (_SZ, [ctypes.POINTER(Batch), _INT]),
    "adrt_hip_queue_plan": (_INT, [ctypes.POINTER(Batch), _INT, _INT, _INT, _INT, _VP, _SZ, _VP,
                                   _SZ, ctypes.POINTER(Queue)]),
    "adrt_hip_queue_run": (_INT, [ctypes.POINTER(Queue), _INT, _VP]),
    "adrt_hip_queue_timeouts": (_INT, [ctypes.POINTER(Queue), ctypes.POINTER(_INT)]),
    "adrt_hip_iadrt": (_INT, [_VP, _VP, _VP, _SZ, _INT, _SZ, _SZ, _VP]),
    "adrt_hip_fmg_restriction": (_INT, [_VP, _VP, _INT, _SZ, _SZ, _VP]),
    "adrt_hip_fmg_prolongation": (_INT, [_VP, _VP, _INT, _SZ, _SZ, _SZ, _VP]),
    "adrt_hip_fmg_highpass": (_INT, [_VP, _VP, _INT, _SZ, _SZ, _SZ, _VP]),
    "adrt_hip_adrt_init": (_INT, [_VP, _VP, _SZ, _SZ, _SZ, _VP]),
    "adrt_hip_truncate": (_INT, [_VP, _VP, _SZ, _SZ, _SZ, _VP]),
    "adrt_hip_truncate_combine": (_INT, [_VP, _VP, _INT, _SZ, _SZ, ctypes.c_double, _INT, _VP]),
    "adrt_hip_stitch": (_INT, [_VP, _VP, _SZ, _SZ, _SZ, _INT, _VP]),
    "adrt_hip_unstitch": (_INT, [_VP, _VP, _SZ, _SZ, _SZ, _INT, _VP]),
    "adrt_hip_interp_table_bytes": (_SZ, [_INT, _SZ]),
    "adrt_hip_interp_table": (_INT, [_INT, _SZ, _VP, _SZ]),
    "adrt_hip_interp_to_cart": (_INT, [_VP, _VP, _VP, _SZ, _INT, _SZ, _SZ, _VP]),
    "adrt_hip_stage_chain_workspace": (_SZ, [_INT, _SZ, _SZ]),
    "adrt_hip_stage_chain": (_INT, [_VP, _VP, _VP, _SZ, _INT, _SZ, _SZ, _INT, _INT, _VP]),
    "adrt_hip_cg_workspace": (_SZ, [_SZ, _SZ]),
    "adrt_hip_batched_dot": (_INT, [_VP, _VP, _VP, _VP, _SZ, _INT, _SZ, _SZ, _VP]),
    "adrt_hip_cg_update": (_INT, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _SZ, _INT, _SZ, _SZ, _VP]),
    "adrt_hip_cg_direction": (_INT, [_VP, _VP, _VP, _INT, _SZ, _SZ, _VP]),
    "adrt_hip_pcg_state_bytes": (_SZ, [_SZ]),
    "adrt_hip_pcg_update": (_INT, [_VP, _VP, _VP, _VP, _VP, _VP, _SZ, _INT, _SZ, _SZ, _VP]),
    "adrt_hip_pcg_direction": (_INT, [_VP, _VP, _VP, _VP, _VP, _SZ, _INT, _SZ, _SZ, _VP]),
    "adrt_hip_set_config": (_INT, [_INT, _SZ, _INT, _SZ]),
    "adrt_hip_get_config": (_INT, [_INT, _SZ, ctypes.POINTER(_INT), ctypes.POINTER(_SZ)]),
    "adrt_hip_nn_conv3x3": (_INT, [_VP, _INT, _VP, _VP, _VP, _INT, _VP, _VP, _VP, _INT, _INT,
                                   _INT, _INT, _INT, _INT, _INT, _VP]),
    "adrt_hip_nn_conv3x3_upcat": (_INT, [_VP, _INT, _INT, _VP, _INT, _VP, _VP, _VP, _INT, _INT,
                                         _INT, _INT, _INT, _INT, _INT, _VP]),
    "adrt_hip_nn_wgrad_workspace": (_SZ, [_INT, _INT, _INT, _INT, _INT]),
    "adrt_hip_nn_wgrad": (_INT, [_VP, _INT, _VP, _INT, _VP, _VP, _VP, _SZ, _INT, _INT, _INT,
                                 _INT, _INT, _VP]),
    "adrt_hip_nn_transpose_weights": (_INT, [_VP, _VP, _INT, _INT, _VP]),
    "adrt_hip_nn_pool": (_INT, [_VP, _INT, _VP, _INT, _INT, _INT, _INT, _INT, _VP]),
    "adrt_hip_nn_pool_backward": (_INT, [_VP, _INT, _VP, _INT, _VP, _INT, _VP, _INT, _INT, _INT,
                                         _INT, _INT, _VP]),
    "adrt_hip_nn_upsample": (_INT, [_VP, _INT, _VP, _INT, _INT, _INT, _INT, _INT, _VP]),
    "adrt_hip_nn_upsample_backward": (_INT, [_VP, _INT, _VP, _INT, _INT, _INT, _INT, _INT, _VP]),
    "adrt_hip_nn_pack_input": (_INT, [_VP, _VP, _INT, _INT, _INT, ctypes.c_float, _VP]),
    "adrt_hip_nn_loss_grad": (_INT, [_VP, _VP, _VP, _INT, _INT, _INT, ctypes.c_float, _VP]),
    "adrt_hip_nn_conv_in": (_INT, [_VP, ctypes.c_float, _VP, _INT, _VP, _VP, _INT, _INT, _INT,
                                   _INT, _VP]),
    "adrt_hip_nn_encoder0": (_INT, [_VP, ctypes.c_float, _VP, _INT, _VP, _VP, _VP, _VP, _INT, _VP,
                                    _INT, _INT, _INT, _INT, _VP]),
    "adrt_hip_nn_conv_out": (_INT, [_VP, _INT, _VP, _VP, _VP, _VP, _INT, _INT, _INT, _VP]),
    "adrt_hip_nn_adam": (_INT, [_VP, _VP, _VP, _VP, _VP, _SZ, ctypes.c_float, ctypes.c_float,
                                ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                _VP]),
    "adrt_hip_pyr_pool": (_INT, [_VP, _VP, _INT, _INT, _INT, _VP]),
    "adrt_hip_pyr_stencil_up": (_INT, [_VP, _VP, _VP, _VP, _INT, _INT, _INT, _VP]),
    "adrt_hip_pyr_correlate": (_INT, [_VP, _VP, _VP, _INT, _INT, _INT, _INT, _VP]),
    "adrt_hip_phantom_raster": (_INT, [_VP, _VP, _INT, _VP, _INT, _INT, _VP]),
    "adrt_hip_dct_table_len": (_SZ, [_SZ]),
    "adrt_hip_dct_table": (_INT, [_SZ, _VP, _SZ]),
    "adrt_hip_dct_filter": (_INT, [_VP, _VP, _VP, _VP, _VP, _SZ, _SZ, _INT, _VP]),
}

_lib = None


def load():
    """Load the library (raises RuntimeError if it was never built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"adrt_amd HIP library not found at {LIB_PATH}; build it with "
                "`python -m adrt_amd.build` (there is no CPU fallback)"
            )
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def shape3(s):
    return _SHAPE3(*[int(v) for v in s])


def shape4(s):
    return _SHAPE4(*[int(v) for v in s])


def check(status, what):
    """Map a non-zero status from the library to an exception."""
    if status == OK:
        return
    lib = load()
    msg = lib.adrt_hip_status_string(status).decode()
    if status == ERR_DEVICE:
        detail = lib.adrt_hip_last_error().decode()
        raise RuntimeError(f"{what}: {msg} ({detail})")
    if status in (ERR_SHAPE, ERR_STEP, ERR_TOO_BIG, ERR_ARG, ERR_SCRATCH):
        raise ValueError(f"{what}: {msg}")
    if status == ERR_DTYPE:
        raise TypeError(f"{what}: {msg}")
    raise RuntimeError(f"{what}: {msg}")
