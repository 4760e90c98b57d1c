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
    return _kernels.adjoint(g.contiguous(), 1.0, _kernels.SUM)


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


def lsq_cg(b, x0=None, *, iters=200, callback=None, check_every=16, graph=True):
    """Least squares min_x 0.5 * ||adrt(x) - b||^2 by conjugate gradients on
    the normal equations A^T A x = A^T b (CGNR: the CGLS iterates in exact
    arithmetic) for a batch of ADRT data b (B, 4, 2n-1, n) (or unbatched).

    Every vector of the iteration lives in image space: one step applies A
    once (q = adrt(p), with the autograd graph) and A^T once (A^T q, the
    autograd backward: the hand-written back-projection kernel), with
    alpha = ||s||^2 / p.(A^T A p) (= ||A p||^2), x += alpha p,
    s -= alpha A^T A p, beta = ||s'||^2 / ||s||^2, p = s' + beta p, in two
    device passes (adrt_hip_pcg_update / adrt_hip_pcg_direction) that keep
    every float64 per-image scalar on the device.  The loss follows the
    recurrence ||r_{k+1}||^2 = ||r_k||^2 - alpha_k ||s_k||^2 (exact in exact
    arithmetic), so no sinogram-sized residual is kept; the first and the
    last history entries are computed directly from b - A x.

    Runs ``iters`` steps; every ``check_every`` steps the host checks whether
    every image's gradient has vanished and stops early if so (an image
    whose gradient vanished meanwhile takes zero steps: alpha = beta = 0).
    With ``graph`` (default) the first block of ``check_every`` steps runs
    eagerly and the following whole blocks replay one captured hipGraph of a
    block (``callback``, if given, runs after every block instead of every
    step).  Returns ``(x, loss_history)`` with iters + 1 per-image float64
    losses (fewer after an early stop)."""
    torch = _torch()
    _check_adrt_data(b)
    dev.require_gpu()
    single = b.dim() == 3
    bb = (b.unsqueeze(0) if single else b).contiguous()
    n = int(bb.shape[-1])
    nb = int(bb.shape[0])
    dt = bb.dtype
    x = torch.zeros(nb, n, n, dtype=dt, device=bb.device) if x0 is None else (
        (x0.unsqueeze(0) if x0.dim() == 2 else x0).to(dt).clone().contiguous())
    r = bb - adrt(x).detach()
    s = _apply_adjoint(x, r).contiguous()
    rr = _bdot(r, r)
    del r
    p = s.clone()
    history = [0.5 * rr.clone()]
    # CG on A^T A x = A^T b with the device-side scalar logic of the PCG
    # kernels (tol 0: every image runs until its gradient vanishes exactly):
    # pcg_update forms p.(A^T A p) = ||A p||^2 in image space
    st = _kernels.PcgState(nb, bb.device)
    st.rz.copy_(_bdot(s, s))
    st.active.copy_((st.rz > 0).int())
    st.iters.zero_()
    st.use_p.zero_()
    ws = _kernels.cg_workspace(x)

    def step():
        pv = p.detach().requires_grad_(True)  # p's storage
        with torch.enable_grad():
            y = adrt(pv)
        (w,) = torch.autograd.grad(y, pv, grad_outputs=y.detach())  # A^T A p
        # alpha = ||s||^2 / p.w, x += a p, s -= a w, ||s'||^2 (one pass)
        _kernels.pcg_update(x, s, p, w.contiguous(), st, ws)
        rr.sub_(st.alpha * st.rz)  # ||r||^2 - alpha ||s||^2
        # beta = ||s'||^2 / ||s||^2, p = s' + beta p
        _kernels.pcg_direction(p, s, None, st, ws)

    blk = max(int(check_every), 1)
    hist_blk = torch.zeros((blk, nb), dtype=torch.float64, device=bb.device)

    def eager_steps(count):
        for _ in range(count):
            step()
            history.append(0.5 * rr.clone())
            if callback is not None:
                callback(x[0] if single else x)

    def vanished():
        return not bool((st.active != 0).any())

    done = 0
    g = None
    while done < iters:
        if vanished():
            break
        if graph and g is None and done >= blk and iters - done >= blk:
            torch.cuda.synchronize(bb.device)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                for j in range(blk):
                    step()
                    hist_blk[j].copy_(rr)
        if g is not None and iters - done >= blk:
            g.replay()
            history.extend(0.5 * h for h in hist_blk.clone())
            if callback is not None:
                callback(x[0] if single else x)
            done += blk
        else:
            k = min(blk, iters - done)
            eager_steps(k)
            done += k
    # the last entry from the residual itself
    r = bb - adrt(x).detach()
    history[-1] = 0.5 * _bdot(r, r)
    return (x[0] if single else x), history
