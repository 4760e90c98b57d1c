"""Learned CG initial guess and preconditioner (adrt_amd.learned, csrc/nn.hip).

The network kernels are checked against a float32 CPU reference of the same
operation (torch on the CPU, fed the same bf16-rounded values); the whole
U-Net's forward and parameter gradients against a CPU autograd model with
the same weights; the multiscale preconditioner against a NumPy float64
restatement; the solver against plain CG; and training against the method
without the network (the network must cut the preconditioned CG iterations
on held-out data)."""

import numpy as np
import pytest

from adrt_amd import learned as LN

gpu = pytest.mark.gpu


# ---------------------------------------------------------------------------
# host-side bookkeeping (no GPU)
# ---------------------------------------------------------------------------
def test_symmetry_classes_cover_stencil():
    E = LN._expansion()
    assert E.shape == (25, 6) and np.all(E.sum(axis=1) == 1)
    # class sizes: centre 1, axis-1 4, diag-1 4, axis-2 4, knight 8, diag-2 4
    assert list(E.sum(axis=0)) == [1, 4, 4, 4, 8, 4]
    k = (E @ np.arange(6)).reshape(5, 5)
    assert np.array_equal(k, k.T) and np.array_equal(k, k[::-1]) and np.array_equal(k, k[:, ::-1])


def test_layer_channels_chain():
    ch = {name: (cin, cout) for name, cin, cout in LN.LAYERS}
    # decoder inputs are skip + upsampled channels
    assert ch["d2a"][0] == ch["e2b"][1] + ch["b3b"][1]
    assert ch["d1a"][0] == ch["e1b"][1] + ch["d2b"][1]
    assert ch["d0a"][0] == ch["e0b"][1] + ch["d1b"][1]
    for a, b in [("e0a", "e0b"), ("e1a", "e1b"), ("e2a", "e2b"), ("b3a", "b3b"),
                 ("d2a", "d2b"), ("d1a", "d1b"), ("d0a", "d0b"), ("d0b", "out")]:
        assert ch[a][1] == ch[b][0]
    assert [ch[n][1] for n in ("e0b", "e1b", "e2b", "b3b")] == [16, 32, 64, 128]
    assert all(cin % 8 == 0 and cout % 16 == 0 for cin, cout in ch.values())


def np_pyramid(r, coef):
    """P r = sum_l D_l^T (h_l * D_l r) in float64 (D: 2x2 mean, D^T: nearest / 4)."""
    levels = coef.shape[0]
    pyr = [np.asarray(r, np.float64)]
    for _ in range(1, levels):
        x = pyr[-1]
        pyr.append(0.25 * ((x[:, ::2, ::2] + x[:, ::2, 1::2]) + (x[:, 1::2, ::2] + x[:, 1::2, 1::2])))
    acc = None
    for lvl in range(levels - 1, -1, -1):
        x = pyr[lvl]
        b, h, w = x.shape
        xp = np.pad(x, ((0, 0), (2, 2), (2, 2)))
        y = np.zeros_like(x)
        k = coef[lvl].reshape(5, 5)
        for dy in range(5):
            for dx in range(5):
                y += k[dy, dx] * xp[:, dy:dy + h, dx:dx + w]
        if acc is not None:
            y += 0.25 * np.repeat(np.repeat(acc, 2, 1), 2, 2)
        acc = y
    return acc


def test_np_pyramid_is_symmetric():
    """The restatement itself: <P x, y> = <x, P y> for symmetric stencils."""
    rng = np.random.default_rng(0)
    coef = rng.standard_normal((5, 6)) @ LN._expansion().T
    x, y = rng.standard_normal((2, 1, 32, 32))
    assert abs(np.vdot(np_pyramid(x, coef), y) - np.vdot(x, np_pyramid(y, coef))) < 1e-9


# ---------------------------------------------------------------------------
# kernels vs CPU references
# ---------------------------------------------------------------------------
def _bf(t):
    import torch

    return t.to(torch.bfloat16).float()


def _nchw_weight(w, cin, cout):
    """[cout][9][cin] -> (cout, cin, 3, 3) (tap = ky * 3 + kx)."""
    return w.reshape(cout, 3, 3, cin).permute(0, 3, 1, 2).contiguous()


def _conv_call(x, xs, w, bias, y, ys, B, h, wd, cin, cout, mode, mask=None, ms=0, y32=None,
               resid=None):
    from adrt_amd import _lib

    LN._check(_lib.load().adrt_hip_nn_conv3x3(LN._p(x), xs, LN._p(w), LN._p(bias), LN._p(y), ys,
                                              LN._p(y32), LN._p(resid), LN._p(mask), ms, B, h, wd,
                                              cin, cout, mode, LN._stream()), "conv")


def _bf16_close(got, want):
    got, want = got.float(), want.float()
    tol = 2.0 ** -7 * want.abs() + 1e-5
    return bool(((got - want).abs() <= tol).all())


SHAPES = [(8, 16), (16, 16), (16, 32), (48, 16), (32, 64), (64, 128), (96, 32), (192, 64),
          (128, 128)]


@gpu
class TestConvKernels:
    @pytest.mark.parametrize("cin,cout", SHAPES)
    @pytest.mark.parametrize("hw", [(16, 16), (8, 8), (4, 32)])
    def test_forward_relu_bias(self, cuda, cin, cout, hw):
        import torch
        import torch.nn.functional as F

        h, wd = hw
        B = 2
        g = torch.Generator().manual_seed(cin * 1000 + cout + h)
        x = _bf(torch.randn(B, h, wd, cin, generator=g))
        w = _bf(torch.randn(cout, 9, cin, generator=g) / np.sqrt(9 * cin))
        bias = torch.randn(cout, generator=g) * 0.1
        want = _bf(F.relu(F.conv2d(x.permute(0, 3, 1, 2), _nchw_weight(w, cin, cout), bias,
                                   padding=1)).permute(0, 2, 3, 1))
        xd = x.to(cuda, torch.bfloat16).contiguous()
        wdv = w.to(cuda, torch.bfloat16).contiguous()
        bd = bias.to(cuda)
        y = torch.full((B, h, wd, cout), float("nan"), device=cuda, dtype=torch.bfloat16)
        _conv_call(xd, cin, wdv, bd, y, cout, B, h, wd, cin, cout, 0)
        assert _bf16_close(y.cpu(), want)

    def test_strided_input_and_output(self, cuda):
        """Channels at an offset inside a wider pixel stride (the skip
        concatenation buffers), untouched channels preserved."""
        import torch
        import torch.nn.functional as F

        B, h, cin, cout = 2, 16, 16, 32
        g = torch.Generator().manual_seed(5)
        buf = _bf(torch.randn(B, h, h, 48, generator=g))
        w = _bf(torch.randn(cout, 9, cin, generator=g) / 12)
        bias = torch.zeros(cout)
        out = torch.full((B, h, h, 96), 7.0, dtype=torch.bfloat16, device=cuda)
        bd = buf.to(cuda, torch.bfloat16)
        _conv_call(bd[..., 16:], 48, w.to(cuda, torch.bfloat16), bias.to(cuda), out[..., 32:], 96,
                   B, h, h, cin, cout, 0)
        want = _bf(F.relu(F.conv2d(buf[..., 16:32].permute(0, 3, 1, 2), _nchw_weight(w, cin, cout),
                                   padding=1)).permute(0, 2, 3, 1))
        o = out.cpu()
        assert _bf16_close(o[..., 32:64], want)
        assert torch.all(o[..., :32] == 7.0) and torch.all(o[..., 64:] == 7.0)

    def test_residual_output(self, cuda):
        import torch
        import torch.nn.functional as F

        B, h, cin = 2, 16, 16
        g = torch.Generator().manual_seed(6)
        x = _bf(torch.randn(B, h, h, cin, generator=g))
        w = torch.zeros(16, 9, cin)
        w[0] = _bf(torch.randn(9, cin, generator=g) / 12)
        bias = torch.zeros(16)
        bias[0] = 0.25
        resid = torch.randn(B, h, h, generator=g)
        y32 = torch.empty(B, h, h, device=cuda)
        _conv_call(x.to(cuda, torch.bfloat16), cin, w.to(cuda, torch.bfloat16), bias.to(cuda), None,
                   0, B, h, h, cin, 16, 1, y32=y32, resid=resid.to(cuda))
        want = resid + F.conv2d(x.permute(0, 3, 1, 2), _nchw_weight(w, cin, 16), bias,
                                padding=1)[:, 0]
        assert torch.allclose(y32.cpu(), want, rtol=1e-5, atol=1e-5)

    @pytest.mark.parametrize("cin,cout", [(16, 16), (48, 16), (96, 32), (192, 64), (64, 128)])
    def test_input_gradient(self, cuda, cin, cout):
        """Transposed weights + mode 2 (mask by the forward input) and mode 3
        give conv2d's input gradient."""
        import torch
        from adrt_amd import _lib

        B, h = 2, 16
        g = torch.Generator().manual_seed(cin + 7 * cout)
        x = _bf(torch.randn(B, h, h, cin, generator=g))
        w = _bf(torch.randn(cout, 9, cin, generator=g) / np.sqrt(9 * cin))
        dz = _bf(torch.randn(B, h, h, cout, generator=g))
        xt = x.permute(0, 3, 1, 2).clone().requires_grad_(True)
        y = torch.nn.functional.conv2d(xt, _nchw_weight(w, cin, cout), padding=1)
        (gx,) = torch.autograd.grad(y, xt, dz.permute(0, 3, 1, 2))
        gx = gx.permute(0, 2, 3, 1)
        wd_ = w.to(cuda, torch.bfloat16).contiguous()
        wt = torch.empty_like(wd_)
        LN._check(_lib.load().adrt_hip_nn_transpose_weights(LN._p(wd_), LN._p(wt), cin, cout,
                                                            LN._stream()), "transpose")
        dzd = dz.to(cuda, torch.bfloat16)
        out = torch.empty(B, h, h, cin, device=cuda, dtype=torch.bfloat16)
        _conv_call(dzd, cout, wt, None, out, cin, B, h, h, cout, cin, 3)
        assert _bf16_close(out.cpu(), _bf(gx))
        xd = x.to(cuda, torch.bfloat16)
        _conv_call(dzd, cout, wt, None, out, cin, B, h, h, cout, cin, 2, mask=xd, ms=cin)
        assert _bf16_close(out.cpu(), _bf(gx * (x > 0)))

    @pytest.mark.parametrize("cin,cout", [(8, 16), (16, 16), (48, 16), (32, 64), (192, 64),
                                          (128, 128)])
    @pytest.mark.parametrize("hw", [16, 8])
    def test_weight_gradient(self, cuda, cin, cout, hw):
        import torch
        from adrt_amd import _lib

        B = 3
        g = torch.Generator().manual_seed(cin * 31 + cout + hw)
        x = _bf(torch.randn(B, hw, hw, cin, generator=g))
        dz = _bf(torch.randn(B, hw, hw, cout, generator=g))
        wt = torch.zeros(cout, cin, 3, 3, requires_grad=True)
        bt = torch.zeros(cout, requires_grad=True)
        y = torch.nn.functional.conv2d(x.permute(0, 3, 1, 2), wt, bt, padding=1)
        gw, gb = torch.autograd.grad(y, (wt, bt), dz.permute(0, 3, 1, 2))
        want_w = gw.permute(0, 2, 3, 1).reshape(cout, 9, cin)  # [co][ky*3+kx][ci]
        lib = _lib.load()
        nbytes = lib.adrt_hip_nn_wgrad_workspace(B, hw, hw, cin, cout)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=cuda)
        dw = torch.full((cout, 9, cin), float("nan"), device=cuda)
        db = torch.full((cout,), float("nan"), device=cuda)
        # device tensors stay bound to names until the kernel has run: a
        # temporary freed after its pointer is taken can be reused by the
        # next allocation before the launch
        xd = x.to(cuda, torch.bfloat16)
        dzd = dz.to(cuda, torch.bfloat16)
        LN._check(lib.adrt_hip_nn_wgrad(LN._p(xd), cin, LN._p(dzd), cout, LN._p(dw), LN._p(db),
                                        LN._p(ws), nbytes, B, hw, hw, cin, cout, LN._stream()),
                  "wgrad")
        torch.cuda.synchronize()
        scale = float(want_w.abs().max())
        assert torch.allclose(dw.cpu(), want_w, rtol=1e-4, atol=1e-5 * scale)
        assert torch.allclose(db.cpu(), gb, rtol=1e-4, atol=1e-4)

    def test_pool_upsample_and_gradients(self, cuda):
        import torch
        import torch.nn.functional as F
        from adrt_amd import _lib

        lib = _lib.load()
        B, h, c = 2, 16, 32
        g = torch.Generator().manual_seed(11)
        x = _bf(torch.randn(B, h, h, c, generator=g))
        xd = x.to(cuda, torch.bfloat16)
        y = torch.empty(B, h // 2, h // 2, c, device=cuda, dtype=torch.bfloat16)
        LN._check(lib.adrt_hip_nn_pool(LN._p(xd), c, LN._p(y), c, B, h, h, c, LN._stream()), "p")
        want = _bf(F.avg_pool2d(x.permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1))
        assert _bf16_close(y.cpu(), want)
        up = torch.empty(B, h, h, c, device=cuda, dtype=torch.bfloat16)
        LN._check(lib.adrt_hip_nn_upsample(LN._p(y), c, LN._p(up), c, B, h // 2, h // 2, c,
                                           LN._stream()), "u")
        assert torch.equal(up.cpu().float(),
                           y.cpu().float().repeat_interleave(2, 1).repeat_interleave(2, 2))
        # upsample backward: 2x2 sums
        dy = _bf(torch.randn(B, h, h, c, generator=g))
        dyd = dy.to(cuda, torch.bfloat16)
        dx = torch.empty(B, h // 2, h // 2, c, device=cuda, dtype=torch.bfloat16)
        LN._check(lib.adrt_hip_nn_upsample_backward(LN._p(dyd), c, LN._p(dx), c, B, h // 2, h // 2,
                                                    c, LN._stream()), "ub")
        assert _bf16_close(dx.cpu(), _bf(4 * F.avg_pool2d(dy.permute(0, 3, 1, 2), 2)
                                         .permute(0, 2, 3, 1)))
        # pool backward with skip add and mask
        dp = _bf(torch.randn(B, h // 2, h // 2, c, generator=g))
        dk = _bf(torch.randn(B, h, h, c, generator=g))
        ds = torch.empty(B, h, h, c, device=cuda, dtype=torch.bfloat16)
        dpd, dkd = dp.to(cuda, torch.bfloat16), dk.to(cuda, torch.bfloat16)
        LN._check(lib.adrt_hip_nn_pool_backward(LN._p(dpd), c, LN._p(dkd), c, LN._p(xd), c,
                                                LN._p(ds), c, B, h, h, c, LN._stream()), "pb")
        want = (0.25 * dp.repeat_interleave(2, 1).repeat_interleave(2, 2) + dk) * (x > 0)
        assert _bf16_close(ds.cpu(), _bf(want))

    def test_adam(self, cuda):
        import torch
        from adrt_amd import _lib

        rng = np.random.default_rng(3)
        w, gr, m, v = (rng.standard_normal(1000).astype(np.float32) for _ in range(4))
        v = np.abs(v)
        t = 5
        lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
        tw, tg, tm, tv = (torch.from_numpy(a.copy()).to(cuda) for a in (w, gr, m, v))
        w16 = torch.empty(1000, dtype=torch.bfloat16, device=cuda)
        LN._check(_lib.load().adrt_hip_nn_adam(LN._p(tw), LN._p(tg), LN._p(tm), LN._p(tv),
                                               LN._p(w16), 1000, lr, b1, b2, eps, 1 - b1 ** t,
                                               1 - b2 ** t, LN._stream()), "adam")
        m2 = b1 * m + (1 - b1) * gr
        v2 = b2 * v + (1 - b2) * gr * gr
        w2 = w - lr * (m2 / (1 - b1 ** t)) / (np.sqrt(v2 / (1 - b2 ** t)) + eps)
        assert np.allclose(tw.cpu().numpy(), w2, rtol=1e-5, atol=1e-7)
        # (1 - beta) is formed in float32 on the device, in float64 here
        assert np.allclose(tm.cpu().numpy(), m2, rtol=1e-5, atol=1e-7)
        assert np.allclose(tv.cpu().numpy(), v2, rtol=1e-5, atol=1e-7)
        assert torch.equal(w16.cpu(), torch.from_numpy(w2).to(torch.bfloat16)) or \
            _bf16_close(w16.cpu(), torch.from_numpy(w2))


# ---------------------------------------------------------------------------
# the whole U-Net against a CPU autograd model
# ---------------------------------------------------------------------------
def cpu_unet(flat, z):
    """Float32 CPU model of the network with the same parameters and the same
    bf16 rounding points as the device activations."""
    import torch
    import torch.nn.functional as F

    params = {}
    off = 0
    for name, cin, cout in LN.LAYERS:
        nw = cout * 9 * cin
        w = _bf(flat[off:off + nw]).reshape(cout, 9, cin)
        params[name] = (_nchw_weight(w, cin, cout), flat[off + nw:off + nw + cout])
        off += nw + cout

    def conv(x, name, relu=True):
        w, b = params[name]
        y = F.conv2d(x, w, b, padding=1)
        return _bf(F.relu(y)) if relu else y

    B, n, _ = z.shape
    x8 = torch.zeros(B, 8, n, n)
    x8[:, 0] = _bf(z)
    up = lambda t: t.repeat_interleave(2, 2).repeat_interleave(2, 3)  # noqa: E731
    s0 = conv(conv(x8, "e0a"), "e0b")
    s1 = conv(conv(_bf(F.avg_pool2d(s0, 2)), "e1a"), "e1b")
    s2 = conv(conv(_bf(F.avg_pool2d(s1, 2)), "e2a"), "e2b")
    b3 = conv(conv(_bf(F.avg_pool2d(s2, 2)), "b3a"), "b3b")
    d2 = conv(conv(torch.cat([s2, up(b3)], 1), "d2a"), "d2b")
    d1 = conv(conv(torch.cat([s1, up(d2)], 1), "d1a"), "d1b")
    d0 = conv(conv(torch.cat([s0, up(d1)], 1), "d0a"), "d0b")
    return z + conv(d0, "out", relu=False)[:, 0]


@gpu
class TestUNet:
    @pytest.mark.parametrize("n", [64, 256])
    def test_forward_and_gradients_match_cpu(self, cuda, n):
        """Forward and every parameter gradient of the device U-Net (batch 2)
        against the CPU autograd model, at 64 x 64 and at the solver's
        256 x 256."""
        import torch

        print(f"unet forward/gradients batch 2 n={n}", flush=True)
        net = LN.UNet(seed=1, device=cuda)
        # a non-trivial output layer so every gradient path carries signal
        rng = np.random.default_rng(2)
        w = net.w.cpu().numpy()
        o0, b0, cin, cout = net.offsets["out"]
        wo = np.zeros((cout, 9, cin), np.float32)
        wo[0] = rng.standard_normal((9, cin)) * 0.2
        w[o0:b0] = wo.reshape(-1)
        for name, _, c in LN.LAYERS:
            _, bb, _, cc = net.offsets[name]
            w[bb:bb + cc] = rng.standard_normal(cc) * 0.05
        w[b0 + 1:b0 + cout] = 0
        net.load_state(w)
        B = 2
        z = torch.from_numpy(rng.standard_normal((B, n, n)).astype(np.float32))
        target = torch.from_numpy(rng.standard_normal((B, n, n)).astype(np.float32))
        y = net.forward(z.to(cuda)).cpu()
        flat = torch.from_numpy(w).requires_grad_(True)
        y_ref = cpu_unet(flat, z)
        rel = float((y - y_ref.detach()).norm() / (y_ref.detach() - z).norm())
        assert rel < 3e-2, rel
        scale = 1.0 / (B * n * n)
        net.backward((scale, target.to(cuda)), B, n)
        loss = 0.5 * scale * ((y_ref - target) ** 2).sum()
        (g_ref,) = torch.autograd.grad(loss, flat)
        g = net.grad.cpu()
        for name, cin, cout in LN.LAYERS:
            w0, b0, _, _ = net.offsets[name]
            for a, b, what in ((w0, b0, "w"), (b0, b0 + cout, "b")):
                gr, gg = g_ref[a:b], g[a:b]
                if float(gr.norm()) == 0:
                    assert float(gg.norm()) == 0
                    continue
                err = float((gg - gr).norm() / gr.norm())
                cos = float((gg * gr).sum() / (gg.norm() * gr.norm()))
                assert err < 6e-2 and cos > 0.998, (name, what, err, cos)

    @pytest.mark.parametrize("n", [32, 64, 256])
    def test_inference_forward_equals_training_forward(self, cuda, n):
        """The inference forward (skip buffers + decoder convolutions that
        read the upsampled coarse features in place) reproduces the training
        forward (materialised concatenations) bit for bit."""
        import torch

        net = LN.UNet(seed=3, device=cuda)
        rng = np.random.default_rng(n)
        w = net.w.cpu().numpy()
        for name, _, c in LN.LAYERS:
            _, bb, _, cc = net.offsets[name]
            w[bb:bb + cc] = rng.standard_normal(cc) * 0.05
        o0, b0, cin, cout = net.offsets["out"]
        w[o0:o0 + cout * 9 * cin] = rng.standard_normal(cout * 9 * cin) * 0.2
        net.load_state(w)
        z = torch.from_numpy(rng.standard_normal((4, n, n)).astype(np.float32)).to(cuda)
        y_train = net.forward(z, train=True).clone()
        y_inf = net.forward(z, train=False).clone()
        assert float((y_train - z).abs().max()) > 0
        assert torch.equal(y_train, y_inf)

    @pytest.mark.parametrize("h,cs,cu,cout", [(64, 16, 32, 16), (128, 32, 64, 32),
                                              (16, 64, 128, 64), (256, 16, 32, 16)])
    def test_upcat_conv_equals_concatenated_conv(self, cuda, h, cs, cu, cout):
        """adrt_hip_nn_conv3x3_upcat over (skip, up) equals the plain
        convolution over the materialised [skip | upsample(up)] bit for bit
        (ReLU + bias mode)."""
        import torch
        from adrt_amd import _lib

        rng = np.random.default_rng(h + cs)
        B, cin = 2, cs + cu
        bf = torch.bfloat16
        skip = torch.from_numpy(rng.standard_normal((B, h, h, cs)).astype(np.float32)).to(bf).to(cuda)
        up = torch.from_numpy(rng.standard_normal((B, h // 2, h // 2, cu)).astype(np.float32)).to(bf).to(cuda)
        wt = torch.from_numpy(rng.standard_normal(cout * 9 * cin).astype(np.float32) * 0.1).to(bf).to(cuda)
        bias = torch.from_numpy(rng.standard_normal(cout).astype(np.float32)).to(cuda)
        cat = torch.cat([skip, up.repeat_interleave(2, 1).repeat_interleave(2, 2)], dim=3).contiguous()
        y0 = torch.full((B, h, h, cout), float("nan"), dtype=bf, device=cuda)
        y1 = torch.full((B, h, h, cout), float("nan"), dtype=bf, device=cuda)
        lib = _lib.load()
        s = torch.cuda.current_stream().cuda_stream
        assert lib.adrt_hip_nn_conv3x3(cat.data_ptr(), cin, wt.data_ptr(), bias.data_ptr(),
                                       y0.data_ptr(), cout, None, None, None, 0, B, h, h, cin,
                                       cout, 0, s) == 0
        assert lib.adrt_hip_nn_conv3x3_upcat(skip.data_ptr(), cs, cs, up.data_ptr(), cu,
                                             wt.data_ptr(), bias.data_ptr(), y1.data_ptr(), cout,
                                             B, h, h, cin, cout, 0, s) == 0
        assert not bool(torch.isnan(y1.float()).any())
        assert torch.equal(y0, y1)

    def test_training_reduces_loss(self, cuda):
        solver = LN.LearnedSolver(n=64, device=cuda)
        solver.fit_preconditioner(count=8)
        hist = solver.train(60, batch=8, pairs=480, lr=2e-3, log_every=1)
        first = np.mean([h for _, h in hist[:5]])
        last = np.mean([h for _, h in hist[-5:]])
        assert last < 0.8 * first, (first, last)


# ---------------------------------------------------------------------------
# preconditioner, phantoms, solver
# ---------------------------------------------------------------------------
@gpu
class TestPreconditioner:
    def test_apply_matches_numpy(self, cuda):
        import torch

        rng = np.random.default_rng(4)
        pyr = LN.Pyramid(levels=6, theta=rng.standard_normal((6, 6)), device=cuda)
        r = rng.standard_normal((3, 64, 64)).astype(np.float32)
        got = pyr(torch.from_numpy(r).to(cuda)).cpu().numpy()
        want = np_pyramid(r, pyr.coef.cpu().numpy().astype(np.float64))
        assert np.allclose(got, want, rtol=1e-4, atol=1e-4 * np.abs(want).max())

    @pytest.mark.parametrize("shape", [(3, 20, 70), (2, 64, 64), (1, 2, 2), (5, 18, 130),
                                       (2, 256, 256)])
    @pytest.mark.parametrize("with_coarse", [False, True])
    def test_stencil_level_matches_numpy(self, cuda, shape, with_coarse):
        """One level of the up-sweep, y = h * x (5x5, zero padding) +
        coarse[parent] / 4, on tiled and partial-tile shapes, against float64
        numpy (float32 rounding of 26 terms)."""
        import torch
        from adrt_amd import _lib

        rng = np.random.default_rng(sum(shape))
        b, h, w = shape
        x = rng.standard_normal(shape).astype(np.float32)
        coef = rng.standard_normal(25).astype(np.float32)
        coarse = rng.standard_normal((b, h // 2, w // 2)).astype(np.float32)
        xd, cd = torch.from_numpy(x).to(cuda), torch.from_numpy(coef).to(cuda)
        kd = torch.from_numpy(coarse).to(cuda)
        y = torch.full_like(xd, float("nan"))
        lib = _lib.load()
        st = lib.adrt_hip_pyr_stencil_up(xd.data_ptr(), kd.data_ptr() if with_coarse else None,
                                         y.data_ptr(), cd.data_ptr(), b, h, w,
                                         torch.cuda.current_stream().cuda_stream)
        assert st == 0
        xp = np.pad(x.astype(np.float64), ((0, 0), (2, 2), (2, 2)))
        want = np.zeros(shape)
        for dy in range(5):
            for dx in range(5):
                want += coef[dy * 5 + dx] * xp[:, dy:dy + h, dx:dx + w]
        if with_coarse:
            want += 0.25 * np.repeat(np.repeat(coarse.astype(np.float64), 2, 1), 2, 2)
        got = y.cpu().numpy()
        assert np.all(np.isfinite(got))
        assert np.allclose(got, want, rtol=1e-5, atol=1e-5 * np.abs(want).max())

    @pytest.mark.parametrize("shape", [(2, 6, 130), (3, 256, 256), (1, 2, 2), (4, 2, 600),
                                       (5, 4, 4), (2, 8, 64), (1000, 2, 8)])
    def test_pool_level_matches_numpy(self, cuda, shape):
        """2 x 2 averaging 0.25 ((a + b) + (c + d)): exact in float32 order."""
        import torch
        from adrt_amd import _lib

        x = np.random.default_rng(shape[2]).standard_normal(shape).astype(np.float32)
        b, h, w = shape
        xd = torch.from_numpy(x).to(cuda)
        y = torch.full((b, h // 2, w // 2), float("nan"), device=cuda)
        st = _lib.load().adrt_hip_pyr_pool(xd.data_ptr(), y.data_ptr(), b, h, w,
                                           torch.cuda.current_stream().cuda_stream)
        assert st == 0
        want = np.float32(0.25) * ((x[:, ::2, ::2] + x[:, ::2, 1::2])
                                   + (x[:, 1::2, ::2] + x[:, 1::2, 1::2]))
        assert y.cpu().numpy().tobytes() == want.tobytes()

    def test_symmetric(self, cuda):
        import torch

        rng = np.random.default_rng(5)
        pyr = LN.Pyramid(levels=7, theta=rng.standard_normal((7, 6)), device=cuda)
        x, y = (torch.from_numpy(rng.standard_normal((1, 128, 128)).astype(np.float32)).to(cuda)
                for _ in range(2))
        a, b = float(LN.bdot(pyr(x), y)), float(LN.bdot(x, pyr(y)))
        assert abs(a - b) <= 1e-4 * max(abs(a), 1.0)

    def test_fit_is_least_squares_optimum(self, cuda):
        """The fitted coefficients minimise sum ||P(M v) - v||^2: no random
        perturbation (or the unpreconditioned scaling) does better."""
        import torch

        n = 32
        solver = LN.LearnedSolver(n=n, device=cuda)
        x = LN.phantom_batch(n, 3, 0, 4, cuda)
        g = torch.Generator(device=cuda)
        g.manual_seed(3)
        vecs = torch.cat([x, torch.randn(x.shape, generator=g, device=cuda)])
        theta = solver.pyramid.fit(vecs)
        w = LN.normal_op(vecs)

        def obj(th):
            solver.pyramid.set_theta(th)
            return float(((solver.pyramid(w) - vecs).double() ** 2).sum())

        best = obj(theta)
        rng = np.random.default_rng(0)
        for _ in range(4):
            pert = theta + 1e-1 * np.abs(theta).max() * rng.standard_normal(theta.shape)
            assert obj(pert) >= best * (1 - 1e-5)
        ident = np.zeros_like(theta)
        ident[0, 0] = float((vecs * w).sum() / (w * w).sum())
        assert obj(ident) > best

    def test_phantom_raster_matches_phantoms(self, cuda):
        from adrt_amd.phantom import phantoms

        got = LN.phantom_batch(128, 7, 3, 4, cuda).cpu().numpy()
        want = phantoms(4, 128, seed=7, start=3, device=cuda).cpu().numpy()
        # cos/sin come from NumPy here and from the device in phantoms():
        # an ulp can move a pixel centre across an ellipse boundary
        assert np.mean(got != want) < 1e-3


@gpu
class TestDirectConvolutions:
    """The first and last layers' vector-ALU kernels against float64
    convolutions of the same bf16-rounded values (zero padding)."""

    @staticmethod
    def _conv_ref(x, w):
        # x (B, H, W, C) float64, w (Cout, 9, C): y[b, i, j, co]
        B, H, W, C = x.shape
        xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
        y = np.zeros((B, H, W, w.shape[0]))
        for t in range(9):
            dy, dx = divmod(t, 3)
            y += np.einsum("bhwc,oc->bhwo", xp[:, dy:dy + H, dx:dx + W], w[:, t, :])
        return y

    @pytest.mark.parametrize("shape", [(2, 256, 256), (3, 32, 48), (1, 20, 70)])
    def test_conv_in(self, cuda, shape):
        import torch
        from adrt_amd import _lib

        rng = np.random.default_rng(shape[1])
        B, H, W = shape
        z = rng.standard_normal(shape).astype(np.float32)
        w = torch.from_numpy(rng.standard_normal(16 * 9 * 8).astype(np.float32) * 0.3).to(torch.bfloat16)
        bias = rng.standard_normal(16).astype(np.float32) * 0.1
        y = torch.full((B, H, W, 16), float("nan"), dtype=torch.bfloat16, device=cuda)
        # device copies held until the kernel has run
        zd, wd, bd = torch.from_numpy(z).to(cuda), w.to(cuda), torch.from_numpy(bias).to(cuda)
        st = _lib.load().adrt_hip_nn_conv_in(zd.data_ptr(), 1.0, wd.data_ptr(), 8, bd.data_ptr(),
                                             y.data_ptr(), 16, B, H, W,
                                             torch.cuda.current_stream().cuda_stream)
        assert st == 0
        torch.cuda.synchronize()
        zb = torch.from_numpy(z).to(torch.bfloat16).double().numpy()[..., None]
        wf = w.double().numpy().reshape(16, 9, 8)[:, :, :1]
        want = np.maximum(self._conv_ref(zb, wf) + bias, 0)
        got = y.float().cpu().numpy()
        # bf16 output rounding (8 bits) dominates the fp32 accumulation error
        assert np.all(np.abs(got - want) <= 2 ** -7 * np.abs(want) + 1e-6)

    @pytest.mark.parametrize("shape", [(2, 256, 256), (3, 32, 48), (1, 20, 70)])
    def test_conv_out(self, cuda, shape):
        import torch
        from adrt_amd import _lib

        rng = np.random.default_rng(shape[2])
        B, H, W = shape
        x = torch.from_numpy(rng.standard_normal((B, H, W, 16)).astype(np.float32)).to(torch.bfloat16)
        w = torch.from_numpy(rng.standard_normal(16 * 9 * 16).astype(np.float32) * 0.3).to(torch.bfloat16)
        bias = rng.standard_normal(16).astype(np.float32)
        resid = rng.standard_normal((B, H, W)).astype(np.float32)
        y = torch.full((B, H, W), float("nan"), device=cuda)
        xd, wd = x.to(cuda), w.to(cuda)
        bd, rd = torch.from_numpy(bias).to(cuda), torch.from_numpy(resid).to(cuda)
        st = _lib.load().adrt_hip_nn_conv_out(xd.data_ptr(), 16, wd.data_ptr(), bd.data_ptr(),
                                              y.data_ptr(), rd.data_ptr(), B, H, W,
                                              torch.cuda.current_stream().cuda_stream)
        assert st == 0
        torch.cuda.synchronize()
        wf = w.double().numpy().reshape(16, 9, 16)[:1]
        want = resid + (self._conv_ref(x.double().numpy(), wf)[..., 0] + bias[0])
        got = y.cpu().numpy()
        assert np.max(np.abs(got - want)) <= 1e-5 * (np.abs(want).max() + 1)


@gpu
class TestDCTFilter:
    @pytest.mark.parametrize("n,b", [(2, 3), (4, 1), (8, 5), (32, 3), (64, 2), (128, 3),
                                     (256, 2), (512, 1)])
    def test_dct2_matches_scipy(self, cuda, n, b):
        """The device 2-D DCT-II (Makhoul's algorithm over in-LDS FFTs)
        against scipy.fft.dctn(norm='ortho') in float64, including batches
        that leave a work-group partly empty (n < 64)."""
        import torch
        from scipy.fft import dctn

        x = np.random.default_rng(n + b).standard_normal((b, n, n)).astype(np.float32)
        f = LN.DCTFilter(n, device=cuda)
        got = f.dct2(torch.from_numpy(x).to(cuda)).cpu().numpy()
        want = dctn(x.astype(np.float64), axes=(1, 2), norm="ortho")
        assert np.max(np.abs(got - want)) <= 2e-6 * np.sqrt(n) * np.max(np.abs(want))

    @pytest.mark.parametrize("n,b", [(2, 2), (16, 3), (64, 4), (256, 3)])
    def test_filter_matches_scipy(self, cuda, n, b):
        """P x = C^T (H .* C x) for a radial profile, against SciPy in float64."""
        import torch
        from scipy.fft import dctn, idctn

        rng = np.random.default_rng(n)
        x = rng.standard_normal((b, n, n)).astype(np.float32)
        f = LN.DCTFilter(n, device=cuda)
        f.set_profile(rng.uniform(0.1, 2.0, f.nbins))
        got = f(torch.from_numpy(x).to(cuda)).cpu().numpy()
        H = f.h[f.bins]
        want = idctn(H * dctn(x.astype(np.float64), axes=(1, 2), norm="ortho"), axes=(1, 2),
                     norm="ortho")
        assert np.max(np.abs(got - want)) <= 4e-6 * np.sqrt(n) * np.max(np.abs(want))

    def test_symmetric_positive(self, cuda):
        import torch

        rng = np.random.default_rng(9)
        f = LN.DCTFilter(128, device=cuda)
        f.set_profile(rng.uniform(0.1, 2.0, f.nbins))
        x, y = (torch.from_numpy(rng.standard_normal((1, 128, 128)).astype(np.float32)).to(cuda)
                for _ in range(2))
        a, c = float(LN.bdot(f(x), y)), float(LN.bdot(x, f(y)))
        assert abs(a - c) <= 1e-5 * max(abs(a), 1.0)
        assert float(LN.bdot(f(x), x)) > 0

    def test_fit_is_least_squares_optimum(self, cuda):
        """The fitted profile minimises sum ||P(M v) - v||^2: random
        perturbations of it, and the pyramid, do no better."""
        import torch

        n = 32
        solver = LN.LearnedSolver(n=n, device=cuda)
        vecs = solver._fit_vectors(3, 4)
        h = solver.dct.fit(vecs)
        w = LN.normal_op(vecs)

        def obj(prof):
            solver.dct.set_profile(prof)
            return float(((solver.dct(w) - vecs).double() ** 2).sum())

        best = obj(h)
        rng = np.random.default_rng(0)
        for _ in range(4):
            assert obj(h * (1 + 0.1 * rng.standard_normal(h.shape))) >= best * (1 - 1e-5)
        solver.pyramid.fit(vecs)
        assert float(((solver.pyramid(w) - vecs).double() ** 2).sum()) >= best * (1 - 1e-3)

    def test_fewer_iterations_than_pyramid(self, cuda):
        """On held-out 64 x 64 systems the fitted DCT filter needs fewer CG
        iterations than the fitted pyramid, and both fewer than plain CG,
        to the same solution."""
        solver = LN.LearnedSolver(n=64, device=cuda)
        solver.fit_preconditioner(count=8)
        _, _, b = TestSolver()._problem(cuda, count=6)
        xp, itp, cp, _ = LN.pcg(LN.normal_op, b, rtol=1e-5)
        xy, ity, cy, _ = LN.pcg(LN.normal_op, b, prec=solver.pyramid, rtol=1e-5)
        xd, itd, cd, _ = LN.pcg(LN.normal_op, b, prec=solver.dct, rtol=1e-5)
        assert bool(cp.all()) and bool(cy.all()) and bool(cd.all())
        assert float(itd.float().mean()) < float(ity.float().mean()) < float(itp.float().mean())
        assert float((xd - xp).norm() / xp.norm()) < 1e-3


@gpu
class TestSolver:
    def _problem(self, cuda, n=64, count=3):
        solver = LN.LearnedSolver(n=n, device=cuda)
        x, b = solver.data(seed=77, start=0, count=count, noise_seed=5)
        return solver, x, b

    def test_plain_pcg_matches_cg_solve(self, cuda):
        from adrt_amd import cg

        _, _, b = self._problem(cuda)
        x1, it1, conv1, _ = LN.pcg(LN.normal_op, b, rtol=1e-5)
        x2, info, _ = cg.cg_solve(cg.ADRTNormalOperator(64), b, rtol=1e-5)
        assert bool(conv1.all()) and int(info.max()) == 0
        assert float((x1 - x2).norm() / x2.norm()) < 1e-4

    def test_preconditioned_converges_faster_to_same_solution(self, cuda):
        solver, _, b = self._problem(cuda)
        solver.fit_preconditioner(count=8)
        xp, itp, cp, _ = LN.pcg(LN.normal_op, b, rtol=1e-6)
        xq, itq, cq, _ = LN.pcg(LN.normal_op, b, prec=solver.pyramid, rtol=1e-6)
        assert bool(cp.all()) and bool(cq.all())
        assert int(itq.max()) < int(itp.min())
        assert float((xq - xp).norm() / xp.norm()) < 1e-3

    def test_indefinite_preconditioner_falls_back(self, cuda):
        _, _, b = self._problem(cuda)
        xp, itp, cp, _ = LN.pcg(LN.normal_op, b, rtol=1e-5)
        xf, itf, cf, _ = LN.pcg(LN.normal_op, b, prec=lambda r: -r, rtol=1e-5)
        assert bool(cf.all())
        assert float((xf - xp).norm() / xp.norm()) < 1e-3

    def test_committed_solver_reaches_1e4_at_256(self, cuda):
        """The committed 256 x 256 solver (U-Net initial guess + learned
        preconditioner) solves held-out noisy systems to relative residual
        1e-4: the true residual ||b - M x|| / ||b||, recomputed from the
        returned x, is below 1e-4 for every image, in fewer iterations than
        plain CG, and the solution agrees with plain CG's."""
        import os

        print("learned pcg 256 x 256, 4 images", flush=True)
        path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            "adrt_amd", "weights", "learned_256.npz")
        solver = LN.LearnedSolver.load(path, device=cuda)
        _, b = solver.data(31337, 0, 4, noise_seed=31337)
        x0 = solver.initial_guess(b, train=False).clone()
        x, it, conv, _ = LN.pcg(LN.normal_op, b, x0, prec=solver.prec, rtol=1e-4)
        xp, itp, convp, _ = LN.pcg(LN.normal_op, b, rtol=1e-4)
        assert bool(conv.all()) and bool(convp.all())
        e = b - LN.normal_op(x)
        rel = (LN.bdot(e, e) / LN.bdot(b, b)).sqrt()
        assert float(rel.max()) < 1e-4, rel
        assert int(it.max()) < int(itp.min())
        assert float((x - xp).norm() / xp.norm()) < 0.05

    def test_network_cuts_iterations(self, cuda):
        """Trained briefly at 64 x 64, the U-Net's initial guess needs fewer
        preconditioned CG iterations on held-out images than the same
        preconditioned CG started from zero (the method without it)."""
        solver = LN.LearnedSolver(n=64, device=cuda)
        solver.fit_preconditioner(count=16)
        solver.train(300, batch=16, pairs=4800, lr=2e-3)
        x, b = solver.data(seed=999, start=0, count=16, noise_seed=999)
        x0 = solver.initial_guess(b).clone()
        _, it_net, c1, _ = LN.pcg(LN.normal_op, b, x0, prec=solver.prec, rtol=1e-4)
        _, it_zero, c2, _ = LN.pcg(LN.normal_op, b, prec=solver.prec, rtol=1e-4)
        assert bool(c1.all()) and bool(c2.all())
        assert float(it_net.float().mean()) < float(it_zero.float().mean())
