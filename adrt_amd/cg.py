"""Conjugate-gradient inversion of the ADRT, entirely on the device.

Restates the reference's documented recipe (docs/examples.cginverse.md:40-67)
for device tensors:

* ``ADRTNormalOperator``: x -> mean_q(truncate(bdrt(adrt(x)))), the normal
  operator A^T A / 4 (A = adrt; A^T y = sum_q truncate(bdrt(y))_q);
* ``ADRTRidgeOperator``: the ridge variant + lambda * x
  (docs/examples.tomography.md:63-71);
* ``iadrt_cg``: CG on M x = mean_q(truncate(bdrt(b))), same arguments, errors
  and result as the recipe.

The CG loop restates ``scipy.sparse.linalg.cg`` (SciPy 1.15, the solver the
recipe calls): the same iteration, the same stopping test
``||r|| < max(atol, rtol * ||b||)`` at the top of each iteration, the same
``(x, info)`` result (info = 0 on convergence, else the iteration count).
Differences, by design: the operator runs in the data dtype on the GPU,
while inner products, norms and the CG scalars are float64 on the device;
and ``cg_solve`` accepts a leading batch of independent systems (each image
keeps its own scalars and stops on its own).  No reference test covers this
recipe, so agreement with the CPU recipe is checked to a tolerance.
"""

import numpy as np

from . import _device as dev
from . import _kernels


def _as_device(a):
    t, restore = dev.to_device(a)
    return t.contiguous(), restore


class ADRTNormalOperator:
    """M x = mean over quadrants of truncate(bdrt(adrt(x))) for device
    tensors x of shape (..., n, n) (docs/examples.cginverse.md:40-56)."""

    def __init__(self, img_size, dtype=None):
        self.img_size = int(img_size)
        self.dtype = dtype

    def __call__(self, x):
        return _kernels.adjoint(_kernels.adrt(x), 1.0, _kernels.MEAN)

    matvec = __call__


class ADRTRidgeOperator(ADRTNormalOperator):
    """(A^T A / 4 + lambda I) x (docs/examples.tomography.md:63-71)."""

    def __init__(self, img_size, dtype=None, ridge_param=1000.0):
        super().__init__(img_size, dtype=dtype)
        self.ridge_param = float(ridge_param)

    def __call__(self, x):
        return super().__call__(x) + self.ridge_param * x

    matvec = __call__


def normal_rhs(b):
    """mean_q(truncate(bdrt(b))): the right-hand side of the recipe for ADRT
    data b of shape (..., 4, 2n-1, n) (docs/examples.cginverse.md:62)."""
    return _kernels.adjoint(b, 1.0, _kernels.MEAN)


def _dot(u, v):
    """Per-image float64 inner product over the last two axes."""
    return (u.double() * v.double()).sum(dim=(-2, -1))


def cg_solve(op, rhs, x0=None, *, rtol=1e-5, atol=0.0, maxiter=None, callback=None):
    """CG on ``op(x) = rhs`` for a batch of images: rhs is a device tensor of
    shape (B, n, n) (or (n, n)).  Returns ``(x, info, iterations)``: info is a
    per-image int64 tensor (0 = converged, else maxiter), iterations the
    number of operator applications made.  Follows scipy.sparse.linalg.cg
    (SciPy 1.15) per image; converged images are frozen while the others
    continue."""
    torch = dev.torch()
    single = rhs.dim() == 2
    b = rhs.unsqueeze(0) if single else rhs
    nb, n = b.shape[0], b.shape[-1]
    bnrm2 = _dot(b, b).sqrt()
    tol = torch.clamp(rtol * bnrm2, min=float(atol))
    if maxiter is None:
        maxiter = n * n * 10
    if x0 is None:
        x = torch.zeros_like(b)
        r = b.clone()
    else:
        x = (x0.unsqueeze(0) if x0.dim() == 2 else x0).to(b.dtype).clone()
        r = b - op(x) if bool(x.any()) else b.clone()
    zero_rhs = bnrm2 == 0
    x[zero_rhs] = b[zero_rhs]
    active = ~zero_rhs
    info = torch.zeros(nb, dtype=torch.int64, device=b.device)
    p = None
    rho_prev = None
    iters = 0
    for iteration in range(maxiter):
        rn = _dot(r, r).sqrt()
        active = active & ~(rn < tol)
        if not bool(active.any()):
            break
        rho_cur = _dot(r, r)
        if iteration > 0:
            beta = torch.where(active, rho_cur / rho_prev, torch.zeros_like(rho_cur))
            p = p * beta.to(b.dtype)[:, None, None] + r
        else:
            p = r.clone()
        q = op(p)
        iters += 1
        alpha = torch.where(active, rho_cur / _dot(p, q), torch.zeros_like(rho_cur))
        a = alpha.to(b.dtype)[:, None, None]
        x = x + a * p
        r = r - a * q
        rho_prev = torch.where(active, rho_cur, torch.ones_like(rho_cur))
        if callback is not None:
            callback(x[0] if single else x)
    else:
        info[active] = maxiter
    return (x[0] if single else x), (info[0] if single else info), iters


def iadrt_cg(b, /, *, op_cls=ADRTNormalOperator, **kwargs):
    """Iterative inverse of the ADRT by CG on the normal equations
    (docs/examples.cginverse.md:59-67): ``b`` is unbatched ADRT data
    (4, 2n-1, n); keyword arguments go to the solver (rtol, atol, maxiter,
    x0, callback).  NumPy input gives a NumPy result, device tensors stay on
    the device.  Raises ValueError for a batch dimension and when CG does not
    converge, as the recipe does."""
    if b.ndim > 3:
        raise ValueError("batch dimension not supported for iadrt_cg")
    img_size = b.shape[-1]
    tb, restore = _as_device(b)
    linop = op_cls(img_size=img_size, dtype=tb.dtype)
    rhs = normal_rhs(tb)
    x0 = kwargs.pop("x0", None)
    if x0 is not None:
        x0, _ = _as_device(x0 if dev.is_tensor(x0) else np.ascontiguousarray(x0))
        x0 = x0.reshape(img_size, img_size)
    x, info, _ = cg_solve(linop, rhs, x0, **kwargs)
    if int(info) != 0:
        raise ValueError(f"convergence failed (cg status {int(info)})")
    return restore(x.reshape(img_size, img_size))
