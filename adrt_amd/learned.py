"""Learned initial guess and preconditioner for CG inversion of noisy ADRT data.

Problem: given d = adrt(x) + noise, solve the normal equations M x = b with
M = mean_q truncate(bdrt(adrt(.))) (the CG recipe's operator) and
b = mean_q truncate(bdrt(d)), to a relative residual ||b - M x|| / ||b||.

Two learned parts, both applied on the GPU by this package's own kernels
(``csrc/nn.hip``):

* ``Pyramid``: a linear multiscale preconditioner
  P r = sum_l D_l^T (h_l * (D_l r)), D_l = l-fold 2x2 averaging, h_l a 5x5
  stencil with the symmetries of the square (6 coefficients per level, 8
  levels).  P is symmetric by construction.  Its 48 coefficients are fitted by
  least squares, min sum_v ||P(M v) - v||^2 over phantoms and white noise v,
  with the gradient/Gram entries reduced on the device.
* ``UNet``: a 4-level U-Net (channels 16/32/64/128, two 3x3 convolutions
  per level, ReLU, 2x2 average pooling, nearest upsampling, skip
  concatenation), BF16 weights and activations with FP32 accumulation on the
  matrix cores.  Input z = P b; output x0 = z + net(z), trained with Adam on
  ||x0 - target||^2 against the clean phantom (``train_learned``'s default;
  the solution of the training pair's normal equations is the other
  target).  Forward and backward are hand-written kernels; the gradients are
  checked against a float32 CPU autograd model in the tests.

``pcg`` runs preconditioned CG from x0 with float64 scalars per image: one
iteration is the operator, the preconditioner and the device kernels that
hold every per-image scalar (adrt_hip_pcg_update / adrt_hip_pcg_direction),
with the host checking convergence one iteration behind (optionally the
iteration is captured as a hipGraph); an image whose preconditioned step is
not a descent step (P not positive definite there) continues as plain CG
from its current iterate (the fallback), and ``evaluate`` reports every
image's convergence and its error against plain CG's solution and the
phantom.
"""

import ctypes
import math

import numpy as np

from . import _device as dev
from . import _kernels
from . import _lib
from .phantom import ellipse_params

# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------


def _torch():
    return dev.torch()


def _p(t):
    return None if t is None else t.data_ptr()


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _check(status, what):
    _lib.check(status, what)


def normal_op(x):
    """M x = mean_q truncate(bdrt(adrt(x))) for (B, n, n) float32 device images."""
    return _kernels.adjoint(_kernels.adrt(x), 1.0, _kernels.MEAN)


def normal_rhs(d):
    return _kernels.adjoint(d, 1.0, _kernels.MEAN)


def bdot(u, v):
    """Per-image float64 inner products of (B, n, n) tensors (device kernel)."""
    return _kernels.batched_dot(u.contiguous(), v.contiguous())


# ---------------------------------------------------------------------------
# data: phantoms on the device, noisy ADRT data
# ---------------------------------------------------------------------------
MAX_ELLIPSES = 20


def phantom_batch(n, seed, start, count, device=None):
    """(count, n, n) float32 phantoms number start..start+count-1 of a seeded
    set (the parameters of adrt_amd.phantom.ellipse_params), rasterised by the
    device kernel."""
    torch = _torch()
    device = device or torch.device("cuda", torch.cuda.current_device())
    params = np.zeros((count, MAX_ELLIPSES, 7), dtype=np.float64)
    counts = np.zeros(count, dtype=np.int32)
    for i in range(count):
        p = ellipse_params(seed, start + i)
        k = p.shape[0]
        counts[i] = k
        params[i, :k, 0] = p[:, 0]
        params[i, :k, 1] = p[:, 1]
        params[i, :k, 2] = np.cos(p[:, 4])
        params[i, :k, 3] = np.sin(p[:, 4])
        params[i, :k, 4] = p[:, 2]
        params[i, :k, 5] = p[:, 3]
        params[i, :k, 6] = p[:, 5]
    tp = torch.from_numpy(params).to(device)
    tc = torch.from_numpy(counts).to(device)
    out = torch.empty((count, n, n), dtype=torch.float32, device=device)
    _check(_lib.load().adrt_hip_phantom_raster(_p(tp), _p(tc), MAX_ELLIPSES, _p(out), count, n,
                                               _stream()), "phantom_raster")
    return out


def noisy_data(x, sigma_frac, generator):
    """d = adrt(x) + N(0, (sigma_frac * max|adrt(x)|)^2), per image."""
    torch = _torch()
    a = _kernels.adrt(x)
    amax = a.abs().amax(dim=(-3, -2, -1), keepdim=True)
    noise = torch.randn(a.shape, generator=generator, device=a.device, dtype=a.dtype)
    return a + sigma_frac * amax * noise


# ---------------------------------------------------------------------------
# multiscale preconditioner
# ---------------------------------------------------------------------------
_CLASSES = [(0, 0), (0, 1), (1, 1), (0, 2), (1, 2), (2, 2)]


def _expansion():
    """E (25 x 6): stencil entry (dy+2)*5 + dx+2 <- symmetry class."""
    e = np.zeros((25, len(_CLASSES)))
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            key = (min(abs(dy), abs(dx)), max(abs(dy), abs(dx)))
            e[(dy + 2) * 5 + dx + 2, _CLASSES.index(key)] = 1.0
    return e


class Pyramid:
    """P r = sum_l D_l^T (h_l * D_l r) on (B, n, n) float32 device images."""

    def __init__(self, levels=8, theta=None, device=None):
        torch = _torch()
        self.levels = levels
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.E = _expansion()
        if theta is None:
            theta = np.zeros((levels, len(_CLASSES)))
            theta[0, 0] = 1.0
        self.set_theta(theta)

    def set_theta(self, theta):
        torch = _torch()
        self.theta = np.asarray(theta, dtype=np.float64).reshape(self.levels, len(_CLASSES))
        coef = self.theta @ self.E.T  # (levels, 25)
        self.coef = torch.from_numpy(coef.astype(np.float32)).to(self.device)

    def _pools(self, r):
        torch = _torch()
        lib = _lib.load()
        out = [r]
        for _ in range(1, self.levels):
            x = out[-1]
            b, h, w = x.shape
            y = torch.empty((b, h // 2, w // 2), dtype=x.dtype, device=x.device)
            _check(lib.adrt_hip_pyr_pool(_p(x), _p(y), b, h, w, _stream()), "pyr_pool")
            out.append(y)
        return out

    def apply(self, r, coef=None):
        torch = _torch()
        lib = _lib.load()
        coef = self.coef if coef is None else coef
        pyr = self._pools(r.contiguous())
        acc = None
        for lvl in range(self.levels - 1, -1, -1):
            x = pyr[lvl]
            b, h, w = x.shape
            y = torch.empty_like(x)
            _check(lib.adrt_hip_pyr_stencil_up(_p(x), _p(acc), _p(y), _p(coef[lvl]), b, h, w,
                                               _stream()), "pyr_stencil_up")
            acc = y
        return acc

    __call__ = apply

    def _correlations(self, e, w, parts=256):
        """(levels, 25) sums over the batch of sum_p (D_l e)[p] (D_l w)[p + o]."""
        torch = _torch()
        lib = _lib.load()
        pe, pw = self._pools(e), self._pools(w)
        out = torch.empty((self.levels, parts, 25), dtype=torch.float32, device=e.device)
        for lvl in range(self.levels):
            b, h, wd = pe[lvl].shape
            _check(lib.adrt_hip_pyr_correlate(_p(pe[lvl]), _p(pw[lvl]), _p(out[lvl]), parts, b, h,
                                              wd, _stream()), "pyr_correlate")
        return out.double().sum(dim=1).cpu().numpy()

    def fit(self, vectors, op=normal_op, ridge=1e-8):
        """Least-squares coefficients for min ||P(op(v)) - v||^2 over the
        batch ``vectors`` (B, n, n).  The gradient of the objective is
        (sum_p D_l e * shifted D_l w) tied over the symmetry classes, with
        e = P w - v, w = op(v); it is affine in the coefficients, so 49
        evaluations give the normal equations exactly."""
        torch = _torch()
        w = op(vectors)
        nparam = self.levels * len(_CLASSES)
        zero = torch.zeros((self.levels, 25), dtype=torch.float32, device=self.device)

        def grad(coef):
            e = self.apply(w, coef) - vectors
            return (self._correlations(e, w) @ self.E).reshape(-1)

        c = -grad(zero)
        G = np.empty((nparam, nparam))
        for k in range(nparam):
            th = np.zeros(nparam)
            th[k] = 1.0
            coef = torch.from_numpy((th.reshape(self.levels, -1) @ self.E.T).astype(np.float32))
            G[:, k] = grad(coef.to(self.device)) + c
        G = 0.5 * (G + G.T)
        G += ridge * np.trace(G) / nparam * np.eye(nparam)
        self.set_theta(np.linalg.solve(G, c))
        return self.theta


class DCTFilter:
    """P r = C^T (H .* C r) on (B, n, n) float32 device images: C the
    orthonormal 2-D DCT-II, H[u][v] = h[floor(sqrt(u^2 + v^2))] a radial
    profile of positive weights (symmetric positive definite P).  The DCT
    diagonalises convolutions with symmetric boundary extension, so a radial
    H acts like a ramp-type filter on the image without the wrap-around of a
    periodic FFT filter.  Applied by adrt_hip_dct_filter (three row passes of
    in-LDS FFTs)."""

    def __init__(self, n, h=None, device=None):
        torch = _torch()
        self.n = int(n)
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        lib = _lib.load()
        m = int(lib.adrt_hip_dct_table_len(self.n))
        tab = np.zeros(2 * m, dtype=np.float32)
        _check(lib.adrt_hip_dct_table(self.n, tab.ctypes.data, tab.size), "dct_table")
        self.table = torch.from_numpy(tab).to(self.device)
        u = np.arange(self.n)
        self.bins = np.floor(np.sqrt(u[:, None] ** 2 + u[None, :] ** 2)).astype(np.int64)
        self.nbins = int(self.bins.max()) + 1
        self.set_profile(np.ones(self.nbins) if h is None else h)

    def set_profile(self, h):
        torch = _torch()
        self.h = np.asarray(h, dtype=np.float64).reshape(self.nbins)
        self.H = torch.from_numpy(self.h[self.bins].astype(np.float32)).to(self.device)

    def _run(self, r, mode):
        torch = _torch()
        r = r.contiguous()
        b, n = r.shape[0], r.shape[-1]
        if n != self.n or r.dtype != torch.float32:
            raise ValueError(f"DCTFilter of side {self.n} applied to {tuple(r.shape)} {r.dtype}")
        y = torch.empty_like(r)
        tmp = torch.empty_like(r)
        _check(_lib.load().adrt_hip_dct_filter(_p(r), _p(y), _p(self.table), _p(self.H), _p(tmp),
                                               b, n, mode, _stream()), "dct_filter")
        return y

    def dct2(self, r):
        """The orthonormal 2-D DCT-II of each image (scipy.fft.dctn norm='ortho')."""
        return self._run(r, 0)

    def apply(self, r):
        return self._run(r, 1)

    __call__ = apply

    def fit(self, vectors, op=normal_op):
        """Least-squares profile for min sum_v ||P(op(v)) - v||^2 over the
        batch ``vectors``.  C is orthonormal, so the objective is
        sum_v sum_(u,w) (h[bin(u, w)] (C op v)[u][w] - (C v)[u][w])^2 and each
        bin's weight has the closed form sum (C op v)(C v) / sum (C op v)^2
        (float64 sums; a bin with no energy keeps weight 0, then the smallest
        positive weight, so that P stays positive definite)."""
        torch = _torch()
        a = self.dct2(op(vectors)).double()
        c = self.dct2(vectors.contiguous()).double()
        num = (a * c).sum(0).cpu().numpy()
        den = (a * a).sum(0).cpu().numpy()
        s = np.bincount(self.bins.ravel(), num.ravel(), minlength=self.nbins)
        t = np.bincount(self.bins.ravel(), den.ravel(), minlength=self.nbins)
        h = np.where(t > 0, s / np.where(t > 0, t, 1.0), 0.0)
        pos = h[h > 0]
        floor = pos.min() if pos.size else 1.0
        self.set_profile(np.where(h > 0, h, floor))
        return self.h


# ---------------------------------------------------------------------------
# U-Net
# ---------------------------------------------------------------------------
# (name, cin, cout): cin/cout as stored (multiples of 8 / 16); the first
# layer's real input is 1 channel (padded to 8), the last layer's real
# output 1 channel (padded to 16).
LAYERS = [
    ("e0a", 8, 16), ("e0b", 16, 16),
    ("e1a", 16, 32), ("e1b", 32, 32),
    ("e2a", 32, 64), ("e2b", 64, 64),
    ("b3a", 64, 128), ("b3b", 128, 128),
    ("d2a", 192, 64), ("d2b", 64, 64),
    ("d1a", 96, 32), ("d1b", 32, 32),
    ("d0a", 48, 16), ("d0b", 16, 16),
    ("out", 16, 16),
]
REAL_CIN = {"e0a": 1}
REAL_COUT = {"out": 1}


class UNet:
    """Parameters (fp32 master + bf16 copy), Adam state, and the per-batch
    activation / gradient buffers of the U-Net at side n."""

    def __init__(self, seed=0, device=None):
        torch = _torch()
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.offsets = {}
        off = 0
        for name, cin, cout in LAYERS:
            nw = cout * 9 * cin
            self.offsets[name] = (off, off + nw, cin, cout)
            off += nw + cout
        self.count = off
        rng = np.random.default_rng(seed)
        flat = np.zeros(self.count, dtype=np.float32)
        for name, cin, cout in LAYERS:
            w0, b0, _, _ = self.offsets[name]
            rcin, rcout = REAL_CIN.get(name, cin), REAL_COUT.get(name, cout)
            std = math.sqrt(2.0 / (9 * rcin)) if name != "out" else 1e-3
            w = np.zeros((cout, 9, cin), dtype=np.float32)
            w[:rcout, :, :rcin] = rng.standard_normal((rcout, 9, rcin)) * std
            flat[w0:w0 + w.size] = w.reshape(-1)
        self.w = torch.from_numpy(flat).to(self.device)
        self.w16 = self.w.to(torch.bfloat16)
        self.grad = torch.zeros_like(self.w)
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)
        self.step_count = 0
        self.wt = {}
        self._buffers = {}
        self.refresh_transposed()

    # -- parameters ----------------------------------------------------------
    def weight(self, name):
        w0, _, cin, cout = self.offsets[name]
        return self.w16[w0:w0 + cout * 9 * cin]

    def bias(self, name):
        _, b0, _, cout = self.offsets[name]
        return self.w[b0:b0 + cout]

    def grads(self, name):
        w0, b0, cin, cout = self.offsets[name]
        return self.grad[w0:w0 + cout * 9 * cin], self.grad[b0:b0 + cout]

    def refresh_transposed(self):
        torch = _torch()
        lib = _lib.load()
        for name, cin, cout in LAYERS:
            if name == "e0a":
                continue
            if name not in self.wt:
                self.wt[name] = torch.empty(cout * 9 * cin, dtype=torch.bfloat16,
                                            device=self.device)
            _check(lib.adrt_hip_nn_transpose_weights(_p(self.weight(name)), _p(self.wt[name]),
                                                     cin, cout, _stream()), "transpose_weights")

    def state_dict(self):
        return {"w": self.w.cpu().numpy(), "m": self.m.cpu().numpy(), "v": self.v.cpu().numpy(),
                "step": self.step_count}

    def load_state(self, w, m=None, v=None, step=0):
        torch = _torch()
        self.w.copy_(torch.as_tensor(np.asarray(w, dtype=np.float32)))
        self.w16.copy_(self.w.to(torch.bfloat16))
        if m is not None:
            self.m.copy_(torch.as_tensor(np.asarray(m, dtype=np.float32)))
            self.v.copy_(torch.as_tensor(np.asarray(v, dtype=np.float32)))
        self.step_count = int(step)
        self.refresh_transposed()

    # -- buffers -------------------------------------------------------------
    def buffers(self, B, n, train=True):
        key = (B, n)
        if key in self._buffers and (not train or "g_out" in self._buffers[key]):
            return self._buffers[key]
        torch = _torch()
        bf = torch.bfloat16
        d = self.device

        def t(hw, c):
            return torch.empty((B, hw, hw, c), dtype=bf, device=d)

        n1, n2, n3 = n // 2, n // 4, n // 8
        buf = {
            "x8": t(n, 8), "t0": t(n, 16), "p0": t(n1, 16),
            "t1": t(n1, 32), "p1": t(n2, 32),
            "t2": t(n2, 64), "p2": t(n3, 64),
            "t3": t(n3, 128), "b3": t(n3, 128),
            "t2d": t(n2, 64), "d2": t(n2, 64), "t1d": t(n1, 32), "d1": t(n1, 32),
            "t0d": t(n, 16), "d0": t(n, 16),
            "y": torch.empty((B, n, n), dtype=torch.float32, device=d),
        }
        if not train:
            self._buffers[key] = buf
            return buf
        buf.update({
            # activations kept for backward: skip connections + upsampled
            # coarse features in one buffer per level
            "cat0": t(n, 48), "cat1": t(n1, 96), "cat2": t(n2, 192),
            # gradients
            "g_out": t(n, 16), "g16a": t(n, 16), "g16b": t(n, 16), "g_cat0": t(n, 48),
            "g32a": t(n1, 32), "g32b": t(n1, 32), "g_cat1": t(n1, 96),
            "g64a": t(n2, 64), "g64b": t(n2, 64), "g_cat2": t(n2, 192),
            "g128a": t(n3, 128), "g128b": t(n3, 128), "g_p2": t(n3, 64),
            "g_p1": t(n2, 32), "g_p0": t(n1, 16),
        })
        lib = _lib.load()
        ws = 0
        for name, cin, cout in LAYERS:
            side = {"e0": n, "d0": n, "ou": n, "e1": n1, "d1": n1, "e2": n2, "d2": n2,
                    "b3": n3}[name[:2]]
            ws = max(ws, int(lib.adrt_hip_nn_wgrad_workspace(B, side, side, cin, cout)))
        buf["wg_ws"] = torch.empty(max(ws, 16), dtype=torch.uint8, device=d)
        self._buffers[key] = buf
        return buf

    # -- kernels -------------------------------------------------------------
    def _conv(self, x, xs, name, y, ys, B, h, mode, mask=None, ms=0, y32=None, resid=None,
              transposed=False):
        _, _, cin, cout = self.offsets[name]
        if transposed:
            w, bias, cin, cout = self.wt[name], None, cout, cin
        else:
            w, bias = self.weight(name), self.bias(name)
        _check(_lib.load().adrt_hip_nn_conv3x3(_p(x), xs, _p(w), _p(bias), _p(y), ys, _p(y32),
                                               _p(resid), _p(mask), ms, B, h, h, cin, cout, mode,
                                               _stream()), f"conv {name}")

    def _wgrad(self, x, xs, dz, dzs, name, B, h, ws):
        _, _, cin, cout = self.offsets[name]
        dw, db = self.grads(name)
        _check(_lib.load().adrt_hip_nn_wgrad(_p(x), xs, _p(dz), dzs, _p(dw), _p(db), _p(ws),
                                             ws.numel(), B, h, h, cin, cout, _stream()),
               f"wgrad {name}")

    def forward(self, z, train=True):
        """x0 = z + net(z) for (B, n, n) float32 device images z; keeps the
        activations for backward() (``train=False`` allocates no gradient
        buffers)."""
        lib = _lib.load()
        B, n, _ = z.shape
        z = z.contiguous()
        f = self.buffers(B, n, train)
        n1, n2, n3 = n // 2, n // 4, n // 8
        s = _stream()
        if not train:
            return self._infer(z, f, B, n)
        cat0, cat1, cat2 = f["cat0"], f["cat1"], f["cat2"]
        # the packed input is kept for e0a's weight gradient
        _check(lib.adrt_hip_nn_pack_input(_p(z), _p(f["x8"]), B, n, n, 1.0, s), "pack_input")
        self._conv_in(z, f["t0"], B, n)
        self._conv(f["t0"], 16, "e0b", cat0, 48, B, n, 0)
        _check(lib.adrt_hip_nn_pool(_p(cat0), 48, _p(f["p0"]), 16, B, n, n, 16, s), "pool")
        self._conv(f["p0"], 16, "e1a", f["t1"], 32, B, n1, 0)
        self._conv(f["t1"], 32, "e1b", cat1, 96, B, n1, 0)
        _check(lib.adrt_hip_nn_pool(_p(cat1), 96, _p(f["p1"]), 32, B, n1, n1, 32, s), "pool")
        self._conv(f["p1"], 32, "e2a", f["t2"], 64, B, n2, 0)
        self._conv(f["t2"], 64, "e2b", cat2, 192, B, n2, 0)
        _check(lib.adrt_hip_nn_pool(_p(cat2), 192, _p(f["p2"]), 64, B, n2, n2, 64, s), "pool")
        self._conv(f["p2"], 64, "b3a", f["t3"], 128, B, n3, 0)
        self._conv(f["t3"], 128, "b3b", f["b3"], 128, B, n3, 0)
        _check(lib.adrt_hip_nn_upsample(_p(f["b3"]), 128, _p(cat2[..., 64:]), 192, B, n3, n3, 128,
                                        s), "upsample")
        self._conv(cat2, 192, "d2a", f["t2d"], 64, B, n2, 0)
        self._conv(f["t2d"], 64, "d2b", f["d2"], 64, B, n2, 0)
        _check(lib.adrt_hip_nn_upsample(_p(f["d2"]), 64, _p(cat1[..., 32:]), 96, B, n2, n2, 64, s),
               "upsample")
        self._conv(cat1, 96, "d1a", f["t1d"], 32, B, n1, 0)
        self._conv(f["t1d"], 32, "d1b", f["d1"], 32, B, n1, 0)
        _check(lib.adrt_hip_nn_upsample(_p(f["d1"]), 32, _p(cat0[..., 16:]), 48, B, n1, n1, 32, s),
               "upsample")
        self._conv(cat0, 48, "d0a", f["t0d"], 16, B, n, 0)
        self._conv(f["t0d"], 16, "d0b", f["d0"], 16, B, n, 0)
        self._conv_out(f["d0"], z, f["y"], B, n)
        return f["y"]

    def _conv_in(self, z, y, B, n):
        """e0a on the vector ALU from the float32 input (one real channel)."""
        _, _, cin, _ = self.offsets["e0a"]
        _check(_lib.load().adrt_hip_nn_conv_in(_p(z), 1.0, _p(self.weight("e0a")), cin,
                                               _p(self.bias("e0a")), _p(y), 16, B, n, n,
                                               _stream()), "conv e0a")

    def _conv_out(self, d0, z, y, B, n):
        """The output layer on the vector ALU: y = z + conv(d0)[channel 0] + bias."""
        _check(_lib.load().adrt_hip_nn_conv_out(_p(d0), 16, _p(self.weight("out")),
                                                _p(self.bias("out")), _p(y), _p(z), B, n, n,
                                                _stream()), "conv out")

    def _upcat(self, f, skip, cs, up, name, y, B, h):
        """Decoder convolution over [skip | nearest-upsampled up], read in
        place (adrt_hip_nn_conv3x3_upcat) where its tiling applies, else
        through a concatenation buffer."""
        _, _, cin, cout = self.offsets[name]
        lib = _lib.load()
        if h % 16 == 0 and (h <= 64 or h % 64 == 0):
            _check(lib.adrt_hip_nn_conv3x3_upcat(
                _p(skip), cs, cs, _p(up), cin - cs, _p(self.weight(name)), _p(self.bias(name)),
                _p(y), cout, B, h, h, cin, cout, 0, _stream()), f"conv {name}")
            return
        torch = _torch()
        key = "cat_" + name
        if key not in f:
            f[key] = torch.empty((B, h, h, cin), dtype=torch.bfloat16, device=self.device)
        cat = f[key]
        cat[..., :cs].copy_(skip)
        _check(lib.adrt_hip_nn_upsample(_p(up), cin - cs, _p(cat[..., cs:]), cin, B, h // 2,
                                        h // 2, cin - cs, _stream()), "upsample")
        self._conv(cat, cin, name, y, cout, B, h, 0)

    def _infer(self, z, f, B, n):
        """Inference form of forward(): the skip connections keep their own
        buffers, each decoder convolution reads the coarse features through
        the upsampling in place, so no concatenation buffer is written, and
        the full-resolution encoder level (e0a, e0b, pooling) is one kernel
        that keeps e0a on chip.  Same arithmetic, bit for bit."""
        torch = _torch()
        lib = _lib.load()
        n1, n2, n3 = n // 2, n // 4, n // 8
        s = _stream()
        for key, hw, c in (("s0", n, 16), ("s1", n1, 32), ("s2", n2, 64)):
            if key not in f:
                f[key] = torch.empty((B, hw, hw, c), dtype=torch.bfloat16, device=self.device)
        if n % 64 == 0:  # e0a, e0b and the pooling in one kernel
            _, _, cin, _ = self.offsets["e0a"]
            _check(lib.adrt_hip_nn_encoder0(_p(z), 1.0, _p(self.weight("e0a")), cin,
                                            _p(self.bias("e0a")), _p(self.weight("e0b")),
                                            _p(self.bias("e0b")), _p(f["s0"]), 16, _p(f["p0"]), 16,
                                            B, n, n, s), "encoder0")
        else:
            self._conv_in(z, f["t0"], B, n)
            self._conv(f["t0"], 16, "e0b", f["s0"], 16, B, n, 0)
            _check(lib.adrt_hip_nn_pool(_p(f["s0"]), 16, _p(f["p0"]), 16, B, n, n, 16, s), "pool")
        self._conv(f["p0"], 16, "e1a", f["t1"], 32, B, n1, 0)
        self._conv(f["t1"], 32, "e1b", f["s1"], 32, B, n1, 0)
        _check(lib.adrt_hip_nn_pool(_p(f["s1"]), 32, _p(f["p1"]), 32, B, n1, n1, 32, s), "pool")
        self._conv(f["p1"], 32, "e2a", f["t2"], 64, B, n2, 0)
        self._conv(f["t2"], 64, "e2b", f["s2"], 64, B, n2, 0)
        _check(lib.adrt_hip_nn_pool(_p(f["s2"]), 64, _p(f["p2"]), 64, B, n2, n2, 64, s), "pool")
        self._conv(f["p2"], 64, "b3a", f["t3"], 128, B, n3, 0)
        self._conv(f["t3"], 128, "b3b", f["b3"], 128, B, n3, 0)
        self._upcat(f, f["s2"], 64, f["b3"], "d2a", f["t2d"], B, n2)
        self._conv(f["t2d"], 64, "d2b", f["d2"], 64, B, n2, 0)
        self._upcat(f, f["s1"], 32, f["d2"], "d1a", f["t1d"], B, n1)
        self._conv(f["t1d"], 32, "d1b", f["d1"], 32, B, n1, 0)
        self._upcat(f, f["s0"], 16, f["d1"], "d0a", f["t0d"], B, n)
        self._conv(f["t0d"], 16, "d0b", f["d0"], 16, B, n, 0)
        self._conv_out(f["d0"], z, f["y"], B, n)
        return f["y"]

    def backward(self, g_scale_target, B, n):
        """Gradients of 0.5 * scale * ||y - target||^2 (scale, target) with
        respect to every parameter, into self.grad (uses the activations of
        the last forward at this (B, n))."""
        scale, target = g_scale_target
        lib = _lib.load()
        f = self.buffers(B, n)
        n1, n2, n3 = n // 2, n // 4, n // 8
        s = _stream()
        ws = f["wg_ws"]
        cat0, cat1, cat2 = f["cat0"], f["cat1"], f["cat2"]
        g_cat0, g_cat1, g_cat2 = f["g_cat0"], f["g_cat1"], f["g_cat2"]
        _check(lib.adrt_hip_nn_loss_grad(_p(f["y"]), _p(target), _p(f["g_out"]), B, n, n,
                                         float(scale), s), "loss_grad")
        # output layer and level-0 decoder
        self._wgrad(f["d0"], 16, f["g_out"], 16, "out", B, n, ws)
        self._conv(f["g_out"], 16, "out", f["g16a"], 16, B, n, 2, mask=f["d0"], ms=16,
                   transposed=True)
        self._wgrad(f["t0d"], 16, f["g16a"], 16, "d0b", B, n, ws)
        self._conv(f["g16a"], 16, "d0b", f["g16b"], 16, B, n, 2, mask=f["t0d"], ms=16,
                   transposed=True)
        self._wgrad(cat0, 48, f["g16b"], 16, "d0a", B, n, ws)
        self._conv(f["g16b"], 16, "d0a", g_cat0, 48, B, n, 2, mask=cat0, ms=48, transposed=True)
        _check(lib.adrt_hip_nn_upsample_backward(_p(g_cat0[..., 16:]), 48, _p(f["g32a"]), 32, B,
                                                 n1, n1, 32, s), "upsample_backward")
        # level-1 decoder
        self._wgrad(f["t1d"], 32, f["g32a"], 32, "d1b", B, n1, ws)
        self._conv(f["g32a"], 32, "d1b", f["g32b"], 32, B, n1, 2, mask=f["t1d"], ms=32,
                   transposed=True)
        self._wgrad(cat1, 96, f["g32b"], 32, "d1a", B, n1, ws)
        self._conv(f["g32b"], 32, "d1a", g_cat1, 96, B, n1, 2, mask=cat1, ms=96, transposed=True)
        _check(lib.adrt_hip_nn_upsample_backward(_p(g_cat1[..., 32:]), 96, _p(f["g64a"]), 64, B,
                                                 n2, n2, 64, s), "upsample_backward")
        # level-2 decoder
        self._wgrad(f["t2d"], 64, f["g64a"], 64, "d2b", B, n2, ws)
        self._conv(f["g64a"], 64, "d2b", f["g64b"], 64, B, n2, 2, mask=f["t2d"], ms=64,
                   transposed=True)
        self._wgrad(cat2, 192, f["g64b"], 64, "d2a", B, n2, ws)
        self._conv(f["g64b"], 64, "d2a", g_cat2, 192, B, n2, 2, mask=cat2, ms=192,
                   transposed=True)
        _check(lib.adrt_hip_nn_upsample_backward(_p(g_cat2[..., 64:]), 192, _p(f["g128a"]), 128,
                                                 B, n3, n3, 128, s), "upsample_backward")
        # bottleneck
        self._wgrad(f["t3"], 128, f["g128a"], 128, "b3b", B, n3, ws)
        self._conv(f["g128a"], 128, "b3b", f["g128b"], 128, B, n3, 2, mask=f["t3"], ms=128,
                   transposed=True)
        self._wgrad(f["p2"], 64, f["g128b"], 128, "b3a", B, n3, ws)
        self._conv(f["g128b"], 128, "b3a", f["g_p2"], 64, B, n3, 3, transposed=True)
        # level-2 encoder
        _check(lib.adrt_hip_nn_pool_backward(_p(f["g_p2"]), 64, _p(g_cat2), 192, _p(cat2), 192,
                                             _p(f["g64a"]), 64, B, n2, n2, 64, s), "pool_backward")
        self._wgrad(f["t2"], 64, f["g64a"], 64, "e2b", B, n2, ws)
        self._conv(f["g64a"], 64, "e2b", f["g64b"], 64, B, n2, 2, mask=f["t2"], ms=64,
                   transposed=True)
        self._wgrad(f["p1"], 32, f["g64b"], 64, "e2a", B, n2, ws)
        self._conv(f["g64b"], 64, "e2a", f["g_p1"], 32, B, n2, 3, transposed=True)
        # level-1 encoder
        _check(lib.adrt_hip_nn_pool_backward(_p(f["g_p1"]), 32, _p(g_cat1), 96, _p(cat1), 96,
                                             _p(f["g32a"]), 32, B, n1, n1, 32, s), "pool_backward")
        self._wgrad(f["t1"], 32, f["g32a"], 32, "e1b", B, n1, ws)
        self._conv(f["g32a"], 32, "e1b", f["g32b"], 32, B, n1, 2, mask=f["t1"], ms=32,
                   transposed=True)
        self._wgrad(f["p0"], 16, f["g32b"], 32, "e1a", B, n1, ws)
        self._conv(f["g32b"], 32, "e1a", f["g_p0"], 16, B, n1, 3, transposed=True)
        # level-0 encoder
        _check(lib.adrt_hip_nn_pool_backward(_p(f["g_p0"]), 16, _p(g_cat0), 48, _p(cat0), 48,
                                             _p(f["g16a"]), 16, B, n, n, 16, s), "pool_backward")
        self._wgrad(f["t0"], 16, f["g16a"], 16, "e0b", B, n, ws)
        self._conv(f["g16a"], 16, "e0b", f["g16b"], 16, B, n, 2, mask=f["t0"], ms=16,
                   transposed=True)
        self._wgrad(f["x8"], 8, f["g16b"], 16, "e0a", B, n, ws)
        return self.grad

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step_count += 1
        t = self.step_count
        _check(_lib.load().adrt_hip_nn_adam(_p(self.w), _p(self.grad), _p(self.m), _p(self.v),
                                            _p(self.w16), self.count, lr, beta1, beta2, eps,
                                            1.0 - beta1 ** t, 1.0 - beta2 ** t, _stream()),
               "adam")
        self.refresh_transposed()

    def loss(self, target):
        y = self._buffers[(target.shape[0], target.shape[-1])]["y"]
        return float(((y.double() - target.double()) ** 2).mean())


# ---------------------------------------------------------------------------
# the learned solver
# ---------------------------------------------------------------------------
class LearnedSolver:
    """U-Net initial guess + learned preconditioners for images of side n:
    ``pyramid`` (the multiscale stencil filter that forms the network's input
    z = P b) and ``dct`` (the radial DCT-domain filter, the preconditioner of
    the CG iterations: measured 7.9 mean iterations from zero on held-out
    256 x 256 systems against the pyramid's 16 and plain CG's 31)."""

    def __init__(self, n=256, sigma_frac=0.005, device=None, seed=0):
        torch = _torch()
        self.n = n
        self.sigma_frac = sigma_frac
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.pyramid = Pyramid(levels=int(math.log2(n)), device=self.device)
        self.dct = DCTFilter(n, device=self.device)
        self.net = UNet(seed=seed, device=self.device)
        # the filter that forms the network's input z = P b
        self.input_prec = "pyramid"

    @property
    def prec(self):
        """The preconditioner of the CG iterations."""
        return self.dct

    def data(self, seed, start, count, noise_seed):
        torch = _torch()
        x = phantom_batch(self.n, seed, start, count, self.device)
        g = torch.Generator(device=self.device)
        g.manual_seed(int(noise_seed))
        d = noisy_data(x, self.sigma_frac, g)
        return x, normal_rhs(d)

    def _fit_vectors(self, seed, count):
        torch = _torch()
        x = phantom_batch(self.n, seed, 0, count, self.device)
        g = torch.Generator(device=self.device)
        g.manual_seed(seed)
        noise = torch.randn(x.shape, generator=g, device=self.device)
        return torch.cat([x, noise])

    def fit_preconditioner(self, seed=1, count=128):
        """Least-squares fits of both preconditioners on ``count`` phantoms
        and ``count`` white-noise images; returns the pyramid's coefficients.
        (With 128 + 128 images the DCT profile needs 8.24 PCG iterations from
        zero on the held-out set, against 8.55 when fitted on 16 + 16.)"""
        vecs = self._fit_vectors(seed, count)
        self.dct.fit(vecs)
        return self.pyramid.fit(vecs)

    def fit_dct(self, seed=1, count=128):
        return self.dct.fit(self._fit_vectors(seed, count))

    def initial_guess(self, b, train=True):
        z = (self.dct if self.input_prec == "dct" else self.pyramid)(b)
        return self.net.forward(z, train)

    def save(self, path, **extra):
        np.savez(path, unet=self.net.w.cpu().numpy(), theta=self.pyramid.theta,
                 dct_h=self.dct.h, input_prec=self.input_prec, n=self.n,
                 sigma_frac=self.sigma_frac, **extra)

    @classmethod
    def load(cls, path, device=None):
        """Weights saved by ``save``; a file without the DCT profile (older
        weights) gets it fitted on load (seed 1, 128 phantoms + 128 noise
        images, the same vectors ``fit_preconditioner`` uses)."""
        with np.load(path, allow_pickle=False) as f:
            solver = cls(n=int(f["n"]), sigma_frac=float(f["sigma_frac"]), device=device)
            solver.pyramid.set_theta(f["theta"])
            solver.net.load_state(f["unet"])
            h = f["dct_h"] if "dct_h" in f.files else None
            solver.input_prec = str(f["input_prec"]) if "input_prec" in f.files else "pyramid"
        if h is None:
            solver.fit_dct()
        else:
            solver.dct.set_profile(h)
        return solver

    def solution_target(self, b, rtol=1e-6, maxiter=200):
        """The normal equations' solution for right-hand sides b (CG with the
        preconditioner), the training target of ``target='solution'``."""
        return pcg(normal_op, b, prec=self.prec, rtol=rtol, maxiter=maxiter, graph=False)[0]

    def train(self, steps, batch=64, seed=1000, pairs=50000, lr=1e-3, log_every=0, log=None,
              lr_fn=None, start=0, target="phantom"):
        """Adam on 0.5 * mean ||x0 - t||^2 over seeded phantom pairs (pair i =
        phantom i of ``seed`` with its batch's noise seed); batches cycle
        through ``pairs`` pairs.  The target t is the clean phantom
        (``target='phantom'``, the default: measured to give the fewest CG
        iterations), the normal equations' solution for the pair's noisy data
        (``target='solution'``: the point CG converges to), or none: with
        ``target='residual'`` the loss is sum_i ||M x0_i - b_i||^2 / ||b_i||^2,
        the quantity CG's stopping test measures (its gradient is
        2 M (M x0 - b) / ||b||^2, M being symmetric); the logged value is then
        the mean relative residual.  Returns the loss history (one value per
        logged step)."""
        hist = []
        n = self.n
        nb = max(pairs // batch, 1)
        for step in range(steps):
            k = (start + step) % nb
            x, b = self.data(seed, k * batch, batch, noise_seed=seed * 100003 + k)
            if target == "residual":
                y = self.initial_guess(b)
                # d/dx0 of sum_i ||M x0_i - b_i||^2 / (batch ||b_i||^2) = 2 M (M x0 - b) / ...
                w = 2.0 / (batch * bdot(b, b))
                g = normal_op((normal_op(y) - b).contiguous()) * w.to(y.dtype)[:, None, None]
                self.net.backward((1.0, (y - g).contiguous()), batch, n)
            else:
                if target == "solution":
                    x = self.solution_target(b)
                self.initial_guess(b)
                self.net.backward((1.0 / (batch * n * n), x), batch, n)
            self.net.adam_step(lr=lr if lr_fn is None else lr_fn(step))
            if log_every and (step % log_every == 0 or step == steps - 1):
                if target == "residual":
                    y = self.net._buffers[(batch, n)]["y"]
                    e = normal_op(y) - b
                    hist.append((step, float((bdot(e, e) / bdot(b, b)).sqrt().mean())))
                else:
                    hist.append((step, self.net.loss(x)))
                if log:
                    log(step, hist[-1][1])
        return hist


def pcg(op, b, x0=None, prec=None, *, rtol=1e-4, maxiter=500, graph=False, compact=0.97):
    """Batched preconditioned CG (plain CG when prec is None) for (B, n, n)
    systems with float64 scalars per image.  Stops each image when
    ||b - M x|| <= rtol * ||b|| (the recursively updated residual).  An image
    whose preconditioned step fails (r.z <= 0 or p.Mp <= 0) continues as plain
    CG from its iterate (steepest-descent restart).

    One iteration is the operator, the preconditioner and the device kernels
    adrt_hip_pcg_update / adrt_hip_pcg_direction, which keep every per-image
    scalar on the device.  The host never waits for the iteration it has just
    queued: it reads the number of running images of the previous iteration
    (copied to pinned memory behind an event) while the GPU works on the
    current one, so the loop ends one masked (no-op) iteration late at most.
    With ``graph`` one iteration is captured as a hipGraph and replayed (off
    by default: on MI355X, 1000 images at 256x256 to 1e-4, it measured
    slower, 158 against 140 ms for plain CG and 67 against 62 ms learned,
    since an iteration is milliseconds of transform work and every
    compaction captures anew).
    When the images still running fall to ``compact`` times the current batch
    (and ``compact`` > 0), they are gathered into a smaller dense batch, so
    converged images cost nothing.  Returns (x, iterations per image,
    converged mask, image-operator applications = sum over iterations of the
    batch size operated on)."""
    torch = _torch()
    B = b.shape[0]
    b = b.contiguous()
    dev_ = b.device
    x = torch.zeros_like(b) if x0 is None else x0.contiguous().clone()
    r = (b - op(x)).contiguous() if x0 is not None else b.clone()
    rr = bdot(r, r)
    st = _kernels.PcgState(B, dev_)
    st.tol.copy_(rtol * bdot(b, b).sqrt())
    if prec is None:
        z, rz = r, rr
        st.use_p.zero_()
    else:
        z = prec(r)
        rz = bdot(r, z)
        bad = rz <= 0
        st.use_p.copy_((~bad).int())
        z = torch.where(bad[:, None, None], r, z)
        rz = torch.where(bad, rr, rz)
    st.rz.copy_(rz)
    p = z.clone()
    st.active.copy_((rr.sqrt() > st.tol).int())
    # results by original index; the working set is a dense sub-batch
    x_out = x
    iters_out = torch.zeros(B, dtype=torch.int64, device=dev_)
    conv_out = torch.zeros(B, dtype=torch.bool, device=dev_)
    idx = torch.arange(B, device=dev_)
    S = {"x": x, "r": r, "p": p, "st": st}

    def make_step(S):
        ws = _kernels.cg_workspace(S["x"])
        xs, rs, ps, ss = S["x"], S["r"], S["p"], S["st"]

        def step():
            q = op(ps)
            _kernels.pcg_update(xs, rs, ps, q, ss, ws)
            _kernels.pcg_direction(ps, rs, None if prec is None else prec(rs), ss, ws)

        if not graph:
            return step
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            step()
        return g.replay

    def write_back(S, rows):
        if S["x"] is not x_out:  # the full batch is written in place
            x_out[rows] = S["x"]
        iters_out[rows] = S["st"].iters.long()
        conv_out[rows] = S["st"].active == 0

    counts = torch.zeros(2, dtype=torch.int32, pin_memory=True)
    events = [torch.cuda.Event(), torch.cuda.Event()]
    stream = torch.cuda.current_stream(dev_)
    # running images after the next-to-last queued iteration (exact here)
    lag = int(st.active.sum())
    prev = None
    run = make_step(S)
    apps = 0
    steps = 0
    while steps < maxiter and lag > 0:
        nb = S["x"].shape[0]
        if compact and nb > 8 and lag <= compact * nb:
            stream.synchronize()
            act = S["st"].active
            exact = int(act.sum())
            if exact == 0:
                break
            write_back(S, idx)
            keep = act.nonzero().squeeze(1)
            idx = idx[keep]
            S = {"x": S["x"][keep].contiguous(), "r": S["r"][keep].contiguous(),
                 "p": S["p"][keep].contiguous(), "st": S["st"].select(keep)}
            run = make_step(S)
            nb, lag, prev = exact, exact, None
        run()
        steps += 1
        apps += nb
        slot = steps & 1
        counts[slot:slot + 1].copy_(S["st"].count, non_blocking=True)
        events[slot].record(stream)
        if prev is not None:
            events[prev].synchronize()
            lag = int(counts[prev])
        prev = slot
    write_back(S, idx)
    return x_out, iters_out, conv_out, apps


def evaluate(solver, count=1000, seed=2000, rtol=1e-4, maxiter=500, repeats=5, compact=0.97):
    """Wall time to relative residual ``rtol`` on ``count`` held-out phantoms
    (phantom seed ``seed``, noise seed ``seed``) for plain CG, CG with the
    multiscale preconditioner from zero, and the learned method (U-Net initial
    guess + preconditioned CG).  Each solve is timed from the right-hand side
    to the converged batch, device synchronised on both sides; the three
    methods are run ``repeats`` times in interleaved rounds (plain, learned,
    preconditioner only, network alone), and each reports the minimum,
    median and maximum of its times.  The speedups are ratios of medians
    (and of minima).  Iteration counts and residuals are those of the last
    round (every round computes the same iterates)."""
    import time

    from . import batched

    torch = _torch()
    x, b = solver.data(seed, 0, count, noise_seed=seed)
    bn = bdot(b, b).sqrt()
    # kernel configuration of the operator's transforms for this batch shape
    tuning = [batched.tune(op, solver.n, count)["choice"]
              for op in (_lib.OP_ADRT, _lib.OP_BDRT)]

    def run(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t0

    def spread(ts):
        ts = sorted(ts)
        return {"seconds_min": round(ts[0], 4), "seconds_median": round(ts[len(ts) // 2], 4),
                "seconds_max": round(ts[-1], 4)}

    def stats(res, ts):
        xs, it, conv, apps = res
        resid = (bdot(b - normal_op(xs), b - normal_op(xs)).sqrt() / bn)
        out = {"seconds": round(sorted(ts)[len(ts) // 2], 4), **spread(ts),
               "iterations_mean": round(float(it.float().mean()), 2),
               "iterations_max": int(it.max()), "image_operator_applications": apps,
               "converged": int(conv.sum()), "max_rel_residual": float(resid.max()),
               "rel_error_vs_phantom": float((bdot(xs - x, xs - x) / bdot(x, x)).sqrt().mean())}
        return out

    # one-time costs (kernel loading, first graph captures, the network's
    # buffers) are paid before any timed solve
    solver.net.buffers(count, solver.n, train=False)
    solver.initial_guess(b, train=False)
    pcg(normal_op, b, rtol=rtol, maxiter=3)
    pcg(normal_op, b, prec=solver.prec, rtol=rtol, maxiter=3)
    times = {"plain": [], "learned": [], "prec": [], "net": []}
    for _ in range(max(int(repeats), 1)):
        plain, t = run(lambda: pcg(normal_op, b, rtol=rtol, maxiter=maxiter, compact=compact))
        times["plain"].append(t)
        learned, t = run(lambda: pcg(normal_op, b, solver.initial_guess(b, train=False),
                                     prec=solver.prec, rtol=rtol, maxiter=maxiter,
                                     compact=compact))
        times["learned"].append(t)
        prec_only, t = run(lambda: pcg(normal_op, b, prec=solver.prec, rtol=rtol,
                                       maxiter=maxiter, compact=compact))
        times["prec"].append(t)
        x0, t = run(lambda: solver.initial_guess(b, train=False))
        times["net"].append(t)
    xs = plain[0]
    # FP64 reference (untimed): plain CG in float64 on the first `nref`
    # systems to a tight tolerance; the learned float32 solutions' relative
    # distance to it
    from . import cg as _cg

    nref = min(64, count)
    b64 = b[:nref].double()
    x64, _, _ = _cg.cg_solve(_cg.ADRTNormalOperator(solver.n), b64, rtol=1e-10, maxiter=2000)
    d = learned[0][:nref].double() - x64
    fp64_err = float((bdot(d, d) / bdot(x64, x64)).sqrt().max())

    def med(k):
        return sorted(times[k])[len(times[k]) // 2]

    out = {
        "count": count, "n": solver.n, "rtol": rtol, "repeats": len(times["plain"]),
        "compact": compact, "network_input": solver.input_prec,
        "tuned_adrt_bdrt": tuning,
        "plain_cg": stats(plain, times["plain"]),
        "preconditioner_only": stats(prec_only, times["prec"]),
        "learned": stats(learned, times["learned"]),
        "network_seconds": round(med("net"), 4),
        "initial_guess_rel_error": float((bdot(x0 - x, x0 - x) / bdot(x, x)).sqrt().mean()),
        "initial_guess_rel_error_vs_solution": float(
            (bdot(x0 - xs, x0 - xs) / bdot(xs, xs)).sqrt().mean()),
        "speedup_learned_vs_plain": round(med("plain") / med("learned"), 3),
        "speedup_learned_vs_plain_min_times": round(min(times["plain"]) / min(times["learned"]),
                                                    3),
        "speedup_preconditioner_only_vs_plain": round(med("plain") / med("prec"), 3),
        "fp64_error": fp64_err,
        "fp64_error_basis": f"max over the first {nref} images of ||x - x64|| / ||x64||, x64 "
                            "= float64 plain CG to relative residual 1e-10",
        "learned_vs_plain_solution_rel_diff": float(
            (bdot(learned[0] - plain[0], learned[0] - plain[0]) / bdot(plain[0], plain[0]))
            .sqrt().max()),
    }
    return out
