"""Differentiable ADRT for torch device tensors, and least squares through it.

* ``ADRTFunction`` (``adrt``): forward = the fused forward ADRT; backward =
  the adjoint A^T g = sum_q truncate(bdrt(g))_q, computed by the fused
  back-projection and the quadrant-sum kernel (the reference's adjoint
  identity, tests/test_bdrt.py:310-325).
* ``BDRTFunction`` (``bdrt``): forward = bdrt; backward = the forward stage
  chain adrt_step(0), ..., adrt_step(L-1) on the ADRT-layout buffer, since
  bdrt_step(k) is the transpose of adrt_step(L-1-k)
  (tests/test_bdrt_step.py:76-92).
* ``lsq_cg``: minimises 0.5 * ||adrt(x) - b||^2 by conjugate gradients on the
  normal equations (CGLS); each A^T application is the autograd backward of
  ``adrt``, i.e. the hand-written back-projection kernel.  Inner products
  and the CG scalars are float64 on the device, per image of a batch.

All of it runs on the GPU; there is no CPU path.
"""

from . import _device as dev
from . import _kernels
from .core import num_iters


def _torch():
    return dev.torch()


# Argument checks in the reference's order (adrt_cdefs_py.cpp:79-154,
# 264-289): array type, ndim, zero dimensions, shape, then device and dtype;
# the messages are the reference's.  Nothing reaches a kernel unchecked, so a
# wrong shape cannot make the library read past a buffer.
def _check_dims(t, lo, hi):
    if not dev.is_tensor(t):
        raise TypeError(f"array must be a torch tensor, but got {type(t).__name__}")
    if not lo <= t.dim() <= hi:
        raise ValueError(f"array must have between {lo} and {hi} dimensions, but had {t.dim()}")
    for i, size in enumerate(t.shape):
        if size == 0:
            raise ValueError(
                f"all array dimensions must be nonzero, but found zero in dimension {i}")


def _check_device_dtype(t):
    if not t.is_cuda:
        raise TypeError("array must be a HIP device tensor, but got a host tensor")
    if dev.dtype_code(t) is None:
        raise TypeError(f"unsupported array dtype {t.dtype}")


def _check_image(x):
    """(n, n) or (B, n, n), n a power of two: the shapes adrt accepts
    (adrt_cdefs_py.cpp:275-289)."""
    _check_dims(x, 2, 3)
    n = int(x.shape[-1])
    if int(x.shape[-2]) != n or n & (n - 1):
        raise ValueError("array must be square with a power of two shape")
    _check_device_dtype(x)


def _check_adrt_data(y):
    """(4, 2n-1, n) or (B, 4, 2n-1, n), n a power of two: the shapes bdrt
    accepts (adrt_cdefs_py.cpp:482-495, adrt_cdefs_common.cpp:185-192)."""
    _check_dims(y, 3, 4)
    n = int(y.shape[-1])
    if int(y.shape[-3]) != 4 or int(y.shape[-2]) != 2 * n - 1 or n & (n - 1):
        raise ValueError("array must have a valid ADRT output shape")
    _check_device_dtype(y)


def _adjoint(g):
    """A^T g = sum over quadrants of truncate(bdrt(g)): (..., 4, 2n-1, n) -> (..., n, n)."""
    return _kernels.truncate_combine(_kernels.bdrt(g.contiguous()), 1.0, _kernels.SUM)


def _stage_chain(h):
    """adrt_step(0) then ... adrt_step(L-1) on an ADRT-layout buffer (one
    library call)."""
    h = h.contiguous()
    if num_iters(int(h.shape[-1])) == 0:
        return h.clone()
    return _kernels.stage_chain(h)


def _function_classes():
    torch = _torch()

    class ADRTFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x):
            return _kernels.adrt(x.contiguous())

        @staticmethod
        def backward(ctx, g):
            return _adjoint(g)

    class BDRTFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, y):
            return _kernels.bdrt(y.contiguous())

        @staticmethod
        def backward(ctx, g):
            return _stage_chain(g)

    return ADRTFunction, BDRTFunction


_classes = None


def _get_classes():
    global _classes
    if _classes is None:
        _classes = _function_classes()
    return _classes


def adrt(x):
    """Differentiable forward ADRT of a device tensor (n, n) or (B, n, n)
    (n a power of two, float32 or float64).  Other shapes raise ValueError
    and other dtypes or host tensors TypeError, before any device work."""
    _check_image(x)
    dev.require_gpu()
    return _get_classes()[0].apply(x)


def bdrt(y):
    """Differentiable back-projection of a device tensor (4, 2n-1, n) or
    (B, 4, 2n-1, n) (n a power of two, float32 or float64).  Other shapes
    raise ValueError and other dtypes or host tensors TypeError."""
    _check_adrt_data(y)
    dev.require_gpu()
    return _get_classes()[1].apply(y)


def _bdot(u, v):
    """Per-image float64 inner product of two (B, ...) tensors (device kernel)."""
    return _kernels.batched_dot(u.contiguous(), v.contiguous())


def _apply_adjoint(x, r):
    """A^T r via the autograd backward of adrt at x (hand-written bdrt)."""
    torch = _torch()
    xv = x.detach().requires_grad_(True)
    with torch.enable_grad():
        y = adrt(xv)
        (g,) = torch.autograd.grad(y, xv, grad_outputs=r)
    return g


def lsq_cg(b, x0=None, *, iters=200, callback=None, check_every=16):
    """Least squares min_x 0.5 * ||adrt(x) - b||^2 by CGLS for a batch of
    ADRT data b (B, 4, 2n-1, n) (or unbatched).  Runs ``iters`` iterations;
    every ``check_every`` iterations the host checks whether every image's
    gradient has vanished and stops early if so (an image whose gradient
    vanished meanwhile takes zero steps: alpha = beta = 0).  Returns
    ``(x, loss_history)``; the history holds per-image losses (float64
    tensors) before the first and after every iteration run.

    Each iteration applies A once and A^T once: the autograd graph of
    ``q = adrt(p)`` also gives ``A^T r`` for the updated residual (the
    backward of a linear map does not depend on the point), so no forward is
    run only to reach its backward."""
    torch = _torch()
    _check_adrt_data(b)
    dev.require_gpu()
    single = b.dim() == 3
    bb = (b.unsqueeze(0) if single else b).contiguous()
    n = int(bb.shape[-1])
    x = torch.zeros(bb.shape[0], n, n, dtype=bb.dtype, device=bb.device) if x0 is None else (
        (x0.unsqueeze(0) if x0.dim() == 2 else x0).to(bb.dtype).clone())
    r = bb - adrt(x).detach()
    s = _apply_adjoint(x, r)
    p = s.clone()
    gamma = _bdot(s, s)
    history = [0.5 * _bdot(r, r)]
    one = torch.ones_like(gamma)
    zero = torch.zeros_like(gamma)
    for k in range(iters):
        if k % max(int(check_every), 1) == 0 and not bool((gamma > 0).any()):
            break
        pv = p.detach().requires_grad_(True)
        with torch.enable_grad():
            y = adrt(pv)
        q = y.detach()
        qq = _bdot(q, q)
        alpha = torch.where(qq > 0, gamma / torch.where(qq > 0, qq, one), zero)
        a = alpha.to(bb.dtype)
        x = x + a[:, None, None] * p
        r = r - a[:, None, None, None] * q
        (s,) = torch.autograd.grad(y, pv, grad_outputs=r)  # A^T r
        gamma_new = _bdot(s, s)
        beta = torch.where(gamma > 0, gamma_new / torch.where(gamma > 0, gamma, one), zero)
        p = s + beta.to(bb.dtype)[:, None, None] * p
        gamma = gamma_new
        history.append(0.5 * _bdot(r, r))
        if callback is not None:
            callback(x[0] if single else x)
    return (x[0] if single else x), history
