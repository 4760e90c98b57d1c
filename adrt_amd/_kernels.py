"""Device-tensor calls into the HIP library (no NumPy, no host copies).

Every function takes contiguous HIP tensors, allocates its result (and any
workspace) from torch's caching allocator on the input's device, launches on
the current stream and returns the result tensor.  The library validates
shapes and returns a status; a failure raises.  The solvers, the autograd
functions and the batched API build on these.
"""

from . import _device as dev
from . import _lib


def _code(t):
    code = dev.dtype_code(t)
    if code is None:
        raise TypeError(f"unsupported array dtype {t.dtype}")
    return code


def _batch_n(t, img_dims):
    b = int(t.shape[0]) if t.dim() == img_dims + 1 else 1
    return b, int(t.shape[-1])


def _ws_call(fn_name, op, t, out):
    lib = _lib.load()
    code = _code(t)
    b, n = _batch_n(t, 2 if op == _lib.OP_ADRT else 3)
    nbytes = lib.adrt_hip_workspace_bytes(op, code, b, n)
    ws = dev.workspace(t, nbytes)
    fn = getattr(lib, fn_name)
    _lib.check(fn(dev.ptr(t), dev.ptr(out), dev.ptr(ws), nbytes, code, b, n,
                  dev.stream_ptr(t)), fn_name)
    return out


def adrt(t):
    n = int(t.shape[-1])
    out = dev.empty_like_dtype(t, (*t.shape[:-2], 4, 2 * n - 1, n))
    return _ws_call("adrt_hip_adrt", _lib.OP_ADRT, t, out)


def bdrt(t):
    return _ws_call("adrt_hip_bdrt", _lib.OP_BDRT, t, dev.empty_like_dtype(t, t.shape))


def bdrt_rows(t, rows):
    """bdrt with rows [0, rows) of every quadrant computed; the rows from
    `rows` on are unspecified (adrt_hip_bdrt_rows)."""
    lib = _lib.load()
    code = _code(t)
    b, n = _batch_n(t, 3)
    out = dev.empty_like_dtype(t, t.shape)
    nbytes = lib.adrt_hip_workspace_bytes(_lib.OP_BDRT, code, b, n)
    ws = dev.workspace(t, nbytes)
    _lib.check(lib.adrt_hip_bdrt_rows(dev.ptr(t), dev.ptr(out), dev.ptr(ws), nbytes, code, b, n,
                                      int(rows), dev.stream_ptr(t)), "adrt_hip_bdrt_rows")
    return out


def iadrt(t):
    return _ws_call("adrt_hip_iadrt", _lib.OP_IADRT, t, dev.empty_like_dtype(t, t.shape))


def _step(fn_name, t, step):
    lib = _lib.load()
    out = dev.empty_like_dtype(t, t.shape)
    b, n = _batch_n(t, 3)
    _lib.check(getattr(lib, fn_name)(dev.ptr(t), dev.ptr(out), _code(t), b, n, int(step),
                                     dev.stream_ptr(t)), fn_name)
    return out


def stage_chain(t, first=0, last=None):
    """adrt_step(first), ..., adrt_step(last - 1) in one call (ADRT layout)."""
    lib = _lib.load()
    b, n = _batch_n(t, 3)
    L = int(lib.adrt_hip_num_iters(n))
    last = L if last is None else last
    out = dev.empty_like_dtype(t, t.shape)
    code = _code(t)
    nbytes = lib.adrt_hip_stage_chain_workspace(code, b, n) if last - first > 1 else 0
    ws = dev.workspace(t, nbytes)
    _lib.check(lib.adrt_hip_stage_chain(dev.ptr(t), dev.ptr(out), dev.ptr(ws), nbytes, code, b, n,
                                        int(first), int(last), dev.stream_ptr(t)), "stage_chain")
    return out


def adrt_step(t, step):
    return _step("adrt_hip_adrt_step", t, step)


def bdrt_step(t, step):
    return _step("adrt_hip_bdrt_step", t, step)


def fmg_restriction(t):
    lib = _lib.load()
    b, n = _batch_n(t, 3)
    out = dev.empty_like_dtype(t, (*t.shape[:-2], n - 1, n // 2))
    _lib.check(lib.adrt_hip_fmg_restriction(dev.ptr(t), dev.ptr(out), _code(t), b, n,
                                            dev.stream_ptr(t)), "fmg_restriction")
    return out


def fmg_prolongation(t):
    lib = _lib.load()
    b = int(t.shape[0]) if t.dim() == 3 else 1
    h, w = int(t.shape[-2]), int(t.shape[-1])
    out = dev.empty_like_dtype(t, (*t.shape[:-2], 2 * h, 2 * w))
    _lib.check(lib.adrt_hip_fmg_prolongation(dev.ptr(t), dev.ptr(out), _code(t), b, h, w,
                                             dev.stream_ptr(t)), "fmg_prolongation")
    return out


def fmg_highpass(t):
    lib = _lib.load()
    b = int(t.shape[0]) if t.dim() == 3 else 1
    h, w = int(t.shape[-2]), int(t.shape[-1])
    out = dev.empty_like_dtype(t, t.shape)
    _lib.check(lib.adrt_hip_fmg_highpass(dev.ptr(t), dev.ptr(out), _code(t), b, h, w,
                                         dev.stream_ptr(t)), "fmg_highpass")
    return out


def adrt_init(t):
    """(..., n, n) any 1/2/4/8-byte dtype -> (..., 4, 2n-1, n)."""
    lib = _lib.load()
    n = int(t.shape[-1])
    b = int(t.shape[0]) if t.dim() == 3 else 1
    out = dev.empty_like_dtype(t, (*t.shape[:-2], 4, 2 * n - 1, n))
    _lib.check(lib.adrt_hip_adrt_init(dev.ptr(t), dev.ptr(out), t.element_size(), b, n,
                                       dev.stream_ptr(t)), "adrt_init")
    return out


def truncate(t):
    """(..., 4, 2n-1, n) any 1/2/4/8-byte dtype -> (..., 4, n, n)."""
    lib = _lib.load()
    n = int(t.shape[-1])
    outer = 1
    for s in t.shape[:-3]:
        outer *= int(s)
    out = dev.empty_like_dtype(t, (*t.shape[:-3], 4, n, n))
    if outer == 0:
        return out
    _lib.check(lib.adrt_hip_truncate(dev.ptr(t), dev.ptr(out), t.element_size(), outer, n,
                                     dev.stream_ptr(t)), "truncate")
    return out


MEAN = 0
SUM = 1


def truncate_combine(t, divisor=1.0, mode=MEAN):
    """combine_q(truncate(t) / divisor): (..., 4, 2n-1, n) -> (..., n, n).

    MEAN reproduces ``np.mean(truncate(t) / divisor, axis=-3)`` bit for bit;
    SUM is the exact adjoint's quadrant sum ((q0+q1)+q2)+q3.
    """
    lib = _lib.load()
    n = int(t.shape[-1])
    b = 1
    for s in t.shape[:-3]:
        b *= int(s)
    out = dev.empty_like_dtype(t, (*t.shape[:-3], n, n))
    _lib.check(lib.adrt_hip_truncate_combine(dev.ptr(t), dev.ptr(out), _code(t), b, n,
                                             float(divisor), int(mode), dev.stream_ptr(t)),
               "truncate_combine")
    return out


def adjoint(t, divisor=1.0, mode=MEAN):
    """truncate_combine(bdrt(t), divisor, mode) with only the rows [0, n) of
    each bdrt quadrant computed (bdrt_rows), the rows truncate keeps."""
    return truncate_combine(bdrt_rows(t, int(t.shape[-1])), divisor, mode)


# ---------------------------------------------------------------------------
# batched CG vector operations: (B, ...) tensors, one system per leading index
# ---------------------------------------------------------------------------

def _bm(t):
    b = int(t.shape[0])
    return b, int(t.numel() // max(b, 1))


def cg_workspace(t):
    b, m = _bm(t)
    return dev.workspace(t, _lib.load().adrt_hip_cg_workspace(b, m))


def batched_dot(u, v, ws=None, out=None):
    """Per-system float64 inner products of two (B, ...) tensors."""
    lib = _lib.load()
    b, m = _bm(u)
    ws = cg_workspace(u) if ws is None else ws
    if out is None:
        out = dev.torch().empty(b, dtype=dev.torch().float64, device=u.device)
    _lib.check(lib.adrt_hip_batched_dot(dev.ptr(u), dev.ptr(v), dev.ptr(out), dev.ptr(ws),
                                        ws.numel(), _code(u), b, m, dev.stream_ptr(u)),
               "batched_dot")
    return out


def cg_update(x, r, p, q, alpha, ws=None, rr=None):
    """In place: x += a p, r -= a q (a = alpha[b] in the element type);
    returns the float64 ||r||^2 per system."""
    lib = _lib.load()
    b, m = _bm(x)
    ws = cg_workspace(x) if ws is None else ws
    if rr is None:
        rr = dev.torch().empty(b, dtype=dev.torch().float64, device=x.device)
    _lib.check(lib.adrt_hip_cg_update(dev.ptr(x), dev.ptr(r), dev.ptr(p), dev.ptr(q),
                                      dev.ptr(alpha), dev.ptr(rr), dev.ptr(ws), ws.numel(),
                                      _code(x), b, m, dev.stream_ptr(x)), "cg_update")
    return rr


def cg_direction(p, z, beta):
    """In place: p = z + c p (c = beta[b] in the element type)."""
    lib = _lib.load()
    b, m = _bm(p)
    _lib.check(lib.adrt_hip_cg_direction(dev.ptr(p), dev.ptr(z), dev.ptr(beta), _code(p), b, m,
                                         dev.stream_ptr(p)), "cg_direction")
    return p


class PcgState:
    """The per-image device state of a batched CG solve (the layout of
    adrt_hip_pcg_state_bytes): float64 fields rz, pq, rr, tol, alpha, beta,
    int32 fields active, next, use_p, iters and the int32 active ``count``,
    each a view into one byte buffer."""

    F64 = ("rz", "pq", "rr", "tol", "alpha", "beta")
    I32 = ("active", "next", "use_p", "iters")

    def __init__(self, batch, device):
        torch = dev.torch()
        self.batch = batch
        nbytes = int(_lib.load().adrt_hip_pcg_state_bytes(batch))
        self.buf = torch.zeros(nbytes, dtype=torch.uint8, device=device)
        for i, k in enumerate(self.F64):
            setattr(self, k, self.buf[8 * batch * i:8 * batch * (i + 1)].view(torch.float64))
        base = 8 * batch * len(self.F64)
        for j, k in enumerate(self.I32):
            setattr(self, k, self.buf[base + 4 * batch * j:base + 4 * batch * (j + 1)]
                    .view(torch.int32))
        off = base + 4 * batch * len(self.I32)
        self.count = self.buf[off:off + 4].view(torch.int32)

    def select(self, keep):
        """A new state holding the images ``keep`` (a device index tensor)."""
        out = PcgState(int(keep.numel()), self.buf.device)
        for k in self.F64 + self.I32:
            getattr(out, k).copy_(getattr(self, k)[keep])
        return out


def pcg_update(x, r, p, q, state, ws):
    """One CG step's update on the device (adrt_hip_pcg_update)."""
    b, m = _bm(x)
    _lib.check(_lib.load().adrt_hip_pcg_update(
        dev.ptr(x), dev.ptr(r), dev.ptr(p), dev.ptr(q), dev.ptr(state.buf), dev.ptr(ws),
        ws.numel(), _code(x), b, m, dev.stream_ptr(x)), "pcg_update")


def pcg_direction(p, r, z, state, ws):
    """The next search direction on the device (adrt_hip_pcg_direction);
    z None for plain CG."""
    b, m = _bm(p)
    _lib.check(_lib.load().adrt_hip_pcg_direction(
        dev.ptr(p), dev.ptr(r), None if z is None else dev.ptr(z), dev.ptr(state.buf),
        dev.ptr(ws), ws.numel(), _code(p), b, m, dev.stream_ptr(p)), "pcg_direction")


_interp_tables = {}


def interp_table(code, n, device):
    """The interpolation table for (dtype, n) on ``device`` (computed on the
    host by the library, uploaded once, cached)."""
    key = (code, n, str(device))
    tab = _interp_tables.get(key)
    if tab is None:
        import numpy as np

        lib = _lib.load()
        nbytes = int(lib.adrt_hip_interp_table_bytes(code, n))
        host = np.zeros(nbytes, dtype=np.uint8)
        _lib.check(lib.adrt_hip_interp_table(code, n, host.ctypes.data, nbytes), "interp_table")
        tab = dev.torch().from_numpy(host).to(device)
        _interp_tables[key] = tab
    return tab


def interp_to_cart(t):
    """(..., 4, 2n-1, n) -> (..., n, 4n), nearest-neighbour interpolation."""
    lib = _lib.load()
    b, n = _batch_n(t, 3)
    code = _code(t)
    tab = interp_table(code, n, t.device)
    out = dev.empty_like_dtype(t, (*t.shape[:-3], n, 4 * n))
    _lib.check(lib.adrt_hip_interp_to_cart(dev.ptr(t), dev.ptr(out), dev.ptr(tab), tab.numel(),
                                           code, b, n, dev.stream_ptr(t)), "interp_to_cart")
    return out
