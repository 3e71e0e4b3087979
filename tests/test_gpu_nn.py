"""GPU: the surrogate's BF16 MFMA kernels against plain FP32 references.

Each HIP kernel is compared with the same operation in float32 torch on the
CPU, on the same bf16-rounded inputs.  Tolerances are written per test: the
kernels accumulate in FP32 and round activations to bf16 (8 significant
bits), so single layers agree to ~1e-2 relative and a whole U-Net backward
pass to a few percent.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def rel(a, b):
    a = a.double()
    b = b.double()
    return float((a - b).norm() / (b.norm() + 1e-30))


@pytest.fixture(scope="module")
def sg(shy):
    from simplehydrology_amd import surrogate
    surrogate.nn()
    return surrogate


def ref_conv(x_nhwc, w_rows, b, taps, cin, cout):
    """x [N,H,W,Cin] f32, w [Cout][taps*Cin] tap-major -> y [N,H,W,Cout]."""
    k = 3 if taps == 9 else 1
    w = w_rows.view(cout, k, k, cin).permute(0, 3, 1, 2)  # OIHW
    y = torch.nn.functional.conv2d(x_nhwc.permute(0, 3, 1, 2), w, b, padding=k // 2)
    return y.permute(0, 2, 3, 1)


@pytest.mark.parametrize("N,H,W,cin,cout,taps", [(2, 16, 24, 8, 32, 9), (1, 32, 32, 64, 96, 9),
                                                 (3, 8, 40, 224, 96, 9), (2, 16, 16, 32, 8, 1),
                                                 (1, 13, 7, 16, 8, 9), (1, 24, 20, 32, 64, 9),
                                                 (2, 8, 16, 32, 64, 1), (1, 5, 9, 96, 128, 9),
                                                 (8, 256, 256, 32, 32, 9)])
def test_conv_forward_dgrad_wgrad(sg, N, H, W, cin, cout, taps):
    g = torch.Generator().manual_seed(cin * 7 + cout)
    x = bf(torch.randn(N, H, W, cin, generator=g))
    w = bf(torch.randn(cout, taps * cin, generator=g) * 0.1)
    b = torch.randn(cout, generator=g) * 0.1
    res = bf(torch.randn(N, H, W, cout, generator=g))
    dev = "cuda:0"
    conv = sg.Conv(cin, cout, taps, np.random.default_rng(0), dev)
    conv.w.copy_(w.to(dev))
    conv.w_bf.copy_(w.to(dev).to(torch.bfloat16))
    conv.b.copy_(b.to(dev))
    xd = x.to(dev).to(torch.bfloat16).contiguous()
    resd = res.to(dev).to(torch.bfloat16).contiguous()
    print("conv", N, H, W, cin, cout, taps, flush=True)
    y = conv.forward(xd, (N, H, W), res=resd, relu=True, out_f32=True)
    yr = torch.relu(ref_conv(x, w, b, taps, cin, cout) + res)
    assert rel(y.cpu(), yr) < 2e-3
    # gradients of a linear (no relu) output against autograd
    gy = bf(torch.randn(N, H, W, cout, generator=g))
    xr = x.clone().requires_grad_(True)
    wr = w.clone().requires_grad_(True)
    br = b.clone().requires_grad_(True)
    out = ref_conv(xr, wr, br, taps, cin, cout)
    out.backward(gy)
    scratch = sg.Scratch(dev)
    dx = conv.backward(xd, gy.to(dev).to(torch.bfloat16).contiguous(), (N, H, W), scratch)
    assert rel(conv.dw.cpu(), wr.grad) < 1e-4  # FP32 accumulation of bf16 products
    assert rel(conv.db.cpu(), br.grad) < 1e-4
    assert rel(dx.float().cpu(), xr.grad) < 1e-2  # dx is rounded to bf16


def test_pool_upsample_channels(sg):
    dev = "cuda:0"
    g = torch.Generator().manual_seed(3)
    N, H, W, C = 2, 8, 12, 16
    x = bf(torch.randn(N, H, W, C, generator=g))
    xd = x.to(dev).to(torch.bfloat16)
    p = torch.empty(N, H // 2, W // 2, C, dtype=torch.bfloat16, device=dev)
    sg._ck(sg.nn().shy_nn_pool2(xd.data_ptr(), p.data_ptr(), N, H, W, C, 0, None))
    pr = x.view(N, H // 2, 2, W // 2, 2, C).mean((2, 4))
    assert rel(p.float().cpu(), pr) < 4e-3
    gx = torch.empty_like(xd)
    sg._ck(sg.nn().shy_nn_pool2(p.data_ptr(), gx.data_ptr(), N, H, W, C, 1, None))
    expect = (0.25 * p.float().cpu()).repeat_interleave(2, 1).repeat_interleave(2, 2)
    assert rel(gx.float().cpu(), expect) < 4e-3
    u = torch.zeros(N, H * 2, W * 2, C + 8, dtype=torch.bfloat16, device=dev)
    sg._ck(sg.nn().shy_nn_up2(xd.data_ptr(), u.data_ptr(), N, H, W, C, C + 8, 0, 0, None))
    ur = x.repeat_interleave(2, 1).repeat_interleave(2, 2)
    assert torch.equal(u[..., :C].float().cpu(), ur)
    back = torch.empty_like(xd)
    sg._ck(sg.nn().shy_nn_up2(u.data_ptr(), back.data_ptr(), N, H, W, C, C + 8, 0, 1, None))
    assert rel(back.float().cpu(), 4 * x) < 4e-3
    s = torch.randn(N, H * 2, W * 2, 8, generator=g).to(torch.bfloat16).to(dev)
    sg._ck(sg.nn().shy_nn_channels(s.data_ptr(), u.data_ptr(), N * H * W * 4, 8, C + 8, C, 0, None))
    assert torch.equal(u[..., C:].cpu(), s.cpu())


def test_adam_closed_form(sg):
    dev = "cuda:0"
    w = torch.tensor([1.0, -2.0, 0.5], device=dev)
    gr = torch.tensor([0.1, -0.3, 0.0], device=dev)
    m = torch.zeros(3, device=dev)
    v = torch.zeros(3, device=dev)
    sg._ck(sg.nn().shy_nn_adam(w.data_ptr(), gr.data_ptr(), m.data_ptr(), v.data_ptr(), None, 3,
                               0.01, 0.9, 0.999, 1e-8, 1, 0.0, None))
    # first Adam step moves each weight by lr * sign(g) (bias-corrected m/sqrt(v))
    exp = torch.tensor([1.0 - 0.01, -2.0 + 0.01, 0.5])
    assert torch.allclose(w.cpu(), exp, atol=1e-6)


@pytest.mark.parametrize("cin,cout,taps", [(32, 64, 9), (8, 32, 9), (96, 8, 1)])
def test_adam_conv_writes_the_flipped_weights(sg, cin, cout, taps):
    """shy_nn_adam_conv: the same update as shy_nn_adam, and wf_bf equal to
    shy_nn_flip_weights of the updated bf16 weights."""
    dev = "cuda:0"
    g = torch.Generator().manual_seed(cin + 3 * cout + taps)
    n = cout * taps * cin
    w0 = torch.randn(n, generator=g)
    gr = torch.randn(n, generator=g) * 0.01
    m0 = torch.randn(n, generator=g).abs() * 0.01
    v0 = torch.randn(n, generator=g).abs() * 1e-4
    bufs = [[t.clone().to(dev) for t in (w0, m0, v0)] for _ in range(2)]
    wb = [torch.empty(n, dtype=torch.bfloat16, device=dev) for _ in range(2)]
    wf = torch.empty(n, dtype=torch.bfloat16, device=dev)
    grd = gr.to(dev)
    sg._ck(sg.nn().shy_nn_adam(sg._p(bufs[0][0]), sg._p(grd), sg._p(bufs[0][1]), sg._p(bufs[0][2]),
                               sg._p(wb[0]), n, 1e-3, 0.9, 0.999, 1e-8, 3, 0.01, None))
    sg._ck(sg.nn().shy_nn_adam_conv(sg._p(bufs[1][0]), sg._p(grd), sg._p(bufs[1][1]),
                                    sg._p(bufs[1][2]), sg._p(wb[1]), sg._p(wf), cin, cout, taps,
                                    1e-3, 0.9, 0.999, 1e-8, 3, 0.01, None))
    for a, b in zip(bufs[0] + [wb[0]], bufs[1] + [wb[1]]):
        assert torch.equal(a.view(torch.int16) if a.dtype == torch.bfloat16 else a.view(torch.int32),
                           b.view(torch.int16) if b.dtype == torch.bfloat16 else b.view(torch.int32))
    wf_ref = torch.empty_like(wf)
    sg._ck(sg.nn().shy_nn_flip_weights(sg._p(wb[0]), sg._p(wf_ref), cin, cout, taps, None))
    assert torch.equal(wf.view(torch.int16), wf_ref.view(torch.int16))
    # and by hand: wf[ci][(taps-1-t)*cout + co] = w[co][t*cin + ci]
    wr = wb[0].cpu().float().view(cout, taps, cin)
    assert torch.equal(wf.cpu().float().view(cin, taps, cout), wr.flip(1).permute(2, 1, 0))


def torch_unet(net, x):
    """The same U-Net in float32 torch on the CPU, sharing net's weights."""
    params = []

    def conv(cv, h, res=None, relu=True):
        w = cv.w.detach().cpu().clone().requires_grad_(True)
        b = cv.b.detach().cpu().clone().requires_grad_(True)
        params.append((cv, w, b))
        y = ref_conv(h, w, b, cv.taps, cv.cin, cv.cout)
        if res is not None:
            y = y + res
        return torch.relu(y) if relu else y

    def stage(s, h):
        a = conv(s.a, h)
        b = conv(s.b, a)
        return conv(s.c, b, res=a)

    skips = []
    h = x
    for k in range(4):
        h = stage(net.e[k], h)
        if k < 3:
            skips.append(h)
            n, hh, ww, c = h.shape
            h = h.view(n, hh // 2, 2, ww // 2, 2, c).mean((2, 4))
    for k in (2, 1, 0):
        u = h.repeat_interleave(2, 1).repeat_interleave(2, 2)
        h = stage(net.d[k], torch.cat([u, skips[k]], -1))
    return conv(net.head, h, relu=False), params


def test_unet_gradients_match_fp32_reference(sg):
    dev = "cuda:0"
    net = sg.UNet(seed=4, device=dev)
    g = torch.Generator().manual_seed(9)
    N, H, W = 2, 32, 32
    x = bf(torch.randn(N, H, W, 8, generator=g) * 0.5)
    t = torch.randn(N, H, W, 3, generator=g) * 0.3
    print("unet forward", flush=True)
    y = net.forward(x.to(dev).to(torch.bfloat16).contiguous())
    loss, gy = sg.mse_loss_and_grad(y, t.to(dev))
    net.backward(gy)
    yr, params = torch_unet(net, x)
    lr = ((yr[..., :3] - t) ** 2).mean()
    lr.backward()
    assert rel(y.cpu()[..., :3], yr.detach()[..., :3]) < 3e-2
    assert abs(loss.item() - lr.item()) < 3e-2 * lr.item()
    worst = 0.0
    for cv, w, b in params:
        worst = max(worst, rel(cv.dw.cpu(), w.grad), rel(cv.db.cpu(), b.grad))
    print("worst relative gradient error", worst, flush=True)
    assert worst < 0.08


@pytest.mark.parametrize("N,H,W,c1,c2,cout", [(2, 16, 24, 64, 32, 32), (1, 32, 16, 96, 64, 64),
                                               (3, 8, 40, 128, 96, 96),
                                               # enough 16x16 tiles for the persistent tile loop
                                               (1, 256, 256, 64, 32, 32), (2, 200, 232, 96, 64, 64)])
def test_conv_upcat_equals_concatenated_input(sg, N, H, W, c1, c2, cout):
    """shy_nn_conv_fwd_upcat (the decoder's first convolution reading the
    upsampled half and the skip straight from their tensors) gives the bits of
    up2 + channel concatenation + the plain convolution, and agrees with an
    FP32 torch convolution of the same input."""
    g = torch.Generator().manual_seed(c1 + c2)
    xh = bf(torch.randn(N, H // 2, W // 2, c1, generator=g))
    sk = bf(torch.randn(N, H, W, c2, generator=g))
    w = bf(torch.randn(cout, 9 * (c1 + c2), generator=g) * 0.05)
    b = torch.randn(cout, generator=g) * 0.1
    dev = "cuda:0"
    conv = sg.Conv(c1 + c2, cout, 9, np.random.default_rng(0), dev)
    conv.w_bf.copy_(w.to(dev).to(torch.bfloat16))
    conv.b.copy_(b.to(dev))
    xhd = xh.to(dev).to(torch.bfloat16).contiguous()
    skd = sk.to(dev).to(torch.bfloat16).contiguous()
    print("upcat", N, H, W, c1, c2, cout, flush=True)
    y = conv.forward_upcat(xhd, skd, (N, H, W))
    u = torch.cat([xhd.repeat_interleave(2, 1).repeat_interleave(2, 2), skd], -1).contiguous()
    y0 = conv.forward(u, (N, H, W))
    assert torch.equal(y.view(torch.int16), y0.view(torch.int16))
    xr = torch.cat([xh.repeat_interleave(2, 1).repeat_interleave(2, 2), sk], -1)
    yr = torch.relu(ref_conv(xr, w, b, 9, c1 + c2, cout))
    assert rel(y.float().cpu(), yr) < 1e-2  # output rounded to bf16


@pytest.mark.parametrize("N,H,W,cin,cout", [(2, 16, 24, 32, 32), (1, 40, 24, 64, 32),
                                             (2, 8, 16, 32, 96),
                                             # enough 16x16 tiles for the persistent tile loop
                                             (1, 256, 256, 32, 64), (2, 200, 232, 64, 96)])
def test_conv_dgrad_epilogue_equals_separate_passes(sg, N, H, W, cin, cout):
    """shy_nn_conv_dgrad with gres and mask gives the bits of the plain data
    gradient followed by shy_nn_add and shy_nn_relu_bwd (the backward pass of
    a residual stage), with either option alone too."""
    g = torch.Generator().manual_seed(cin + 3 * cout)
    dev = "cuda:0"
    conv = sg.Conv(cin, cout, 9, np.random.default_rng(1), dev)
    x = torch.randn(N, H, W, cin, generator=g).to(dev).to(torch.bfloat16).contiguous()
    gy = torch.randn(N, H, W, cout, generator=g).to(dev).to(torch.bfloat16).contiguous()
    gres = torch.randn(N, H, W, cin, generator=g).to(dev).to(torch.bfloat16).contiguous()
    scratch = sg.Scratch(dev)
    shape = (N, H, W)
    print("dgrad epilogue", N, H, W, cin, cout, flush=True)
    d0 = conv.backward(x, gy, shape, scratch)
    both = conv.backward(x, gy, shape, scratch, gres=gres, mask=x)
    assert torch.equal(both.view(torch.int16), sg._relu_bwd(x, sg._add(gres, d0)).view(torch.int16))
    only_mask = conv.backward(x, gy, shape, scratch, mask=x)
    assert torch.equal(only_mask.view(torch.int16), sg._relu_bwd(x, d0).view(torch.int16))
    only_res = conv.backward(x, gy, shape, scratch, gres=gres)
    assert torch.equal(only_res.view(torch.int16), sg._add(gres, d0).view(torch.int16))
    assert (both.float() == 0).float().mean() > 0.3  # the mask did zero about half


@pytest.mark.parametrize("N,H,W,cin,hc,with_res", [(2, 20, 36, 32, 8, True), (1, 16, 48, 64, 4, False),
                                                   (1, 13, 21, 32, 16, True)])
def test_conv_head_equals_conv_then_head(sg, N, H, W, cin, hc, with_res):
    """shy_nn_conv_fwd_head: the 32-channel output is shy_nn_conv_fwd's bit
    for bit, and the head applied to it matches the separate 1x1
    convolution (both are FP32 sums of the same BF16 products, in different
    orders), partial tiles included; without y the head is the same."""
    g = torch.Generator().manual_seed(17 * cin + hc)
    dev = "cuda:0"
    conv = sg.Conv(cin, 32, 9, np.random.default_rng(4), dev)
    head = sg.Conv(32, hc, 1, np.random.default_rng(5), dev)
    conv.b.copy_((torch.randn(32, generator=g) * 0.1).to(dev))
    head.b.copy_((torch.randn(hc, generator=g) * 0.1).to(dev))
    x = torch.randn(N, H, W, cin, generator=g).to(dev).to(torch.bfloat16).contiguous()
    res = (torch.randn(N, H, W, 32, generator=g).to(dev).to(torch.bfloat16).contiguous()
           if with_res else None)
    print("conv+head", N, H, W, cin, hc, with_res, flush=True)
    y0 = conv.forward(x, (N, H, W), res=res)
    o0 = head.forward(y0, (N, H, W), relu=False, out_f32=True)
    y1, o1 = conv.forward_head(x, (N, H, W), head, res=res)
    assert torch.equal(y1.view(torch.int16), y0.view(torch.int16))
    assert rel(o1.cpu(), o0.cpu()) < 1e-6
    # and against float32 on the CPU from the rounded y
    ref = y0.float().cpu() @ head.w_bf.float().cpu().view(hc, 32).t() + head.b.cpu()
    assert rel(o1.cpu(), ref) < 1e-6
    _, o2 = conv.forward_head(x, (N, H, W), head, res=res, keep=False)
    assert torch.equal(o2.view(torch.int32), o1.view(torch.int32))


@pytest.mark.parametrize("N,H,W,cin,cout,with_res", [
    (1, 256, 256, 32, 32, True),    # the tile loop
    (2, 200, 232, 32, 64, False),   # the tile loop, partial tiles, two channel blocks
    (1, 32, 48, 64, 96, True),      # too few tiles: convolution, then pool
])
def test_conv_pool_equals_conv_then_pool(sg, N, H, W, cin, cout, with_res):
    """shy_nn_conv_fwd_pool: the output is shy_nn_conv_fwd's bit for bit and
    the pooled output shy_nn_pool2's on it, bit for bit; the pool also
    against float32 on the CPU from the rounded output."""
    g = torch.Generator().manual_seed(N * H + cout)
    dev = "cuda:0"
    conv = sg.Conv(cin, cout, 9, np.random.default_rng(6), dev)
    conv.b.copy_((torch.randn(cout, generator=g) * 0.1).to(dev))
    x = torch.randn(N, H, W, cin, generator=g).to(dev).to(torch.bfloat16).contiguous()
    res = (torch.randn(N, H, W, cout, generator=g).to(dev).to(torch.bfloat16).contiguous()
           if with_res else None)
    print("conv+pool", N, H, W, cin, cout, with_res, flush=True)
    y0 = conv.forward(x, (N, H, W), res=res)
    p0 = torch.empty(N, H // 2, W // 2, cout, dtype=torch.bfloat16, device=dev)
    assert sg.nn().shy_nn_pool2(sg._p(y0), sg._p(p0), N, H, W, cout, 0, sg._stream()) == 0
    y1, p1 = conv.forward_pool(x, (N, H, W), res=res)
    assert torch.equal(y1.view(torch.int16), y0.view(torch.int16))
    assert torch.equal(p1.view(torch.int16), p0.view(torch.int16))
    yc = y0.float().cpu().view(N, H // 2, 2, W // 2, 2, cout)
    ref = (0.25 * (yc[:, :, 0, :, 0] + yc[:, :, 0, :, 1] + yc[:, :, 1, :, 0]
                   + yc[:, :, 1, :, 1])).to(torch.bfloat16)
    assert torch.equal(p1.cpu().view(torch.int16), ref.view(torch.int16))


def test_conv_rejects_unaligned_bias(sg):
    """The convolution entry points read a lane's four biases in one 16-byte
    load, so they reject a bias that is not 16-byte aligned (before any
    launch)."""
    dev = "cuda:0"
    x = torch.zeros(1, 16, 16, 32, dtype=torch.bfloat16, device=dev)
    w = torch.zeros(32, 9 * 32, dtype=torch.bfloat16, device=dev)
    b = torch.zeros(33, dtype=torch.float32, device=dev)
    y = torch.empty(1, 16, 16, 32, dtype=torch.bfloat16, device=dev)
    rc = sg.nn().shy_nn_conv_fwd(sg._p(x), sg._p(w), b.data_ptr() + 4, None, sg._p(y), None,
                                 1, 16, 16, 32, 32, 9, 1, sg._stream())
    assert rc != 0 and b"aligned" in sg.nn().shy_nn_last_error()
    assert sg.nn().shy_nn_conv_fwd(sg._p(x), sg._p(w), b.data_ptr(), None, sg._p(y), None,
                                   1, 16, 16, 32, 32, 9, 1, sg._stream()) == 0
    # the fused head also reads its weights and biases 16 bytes at a time
    hw = torch.zeros(8 * 32 + 8, dtype=torch.bfloat16, device=dev)
    hb = torch.zeros(12, dtype=torch.float32, device=dev)
    hy = torch.empty(1, 16, 16, 8, dtype=torch.float32, device=dev)
    for wofs, bofs in ((2, 0), (0, 4)):
        rc = sg.nn().shy_nn_conv_fwd_head(sg._p(x), sg._p(w), b.data_ptr(), None, sg._p(y),
                                          hw.data_ptr() + wofs, hb.data_ptr() + bofs,
                                          sg._p(hy), 1, 16, 16, 32, 8, 1, sg._stream())
        assert rc != 0 and b"aligned" in sg.nn().shy_nn_last_error()
    assert sg.nn().shy_nn_conv_fwd_head(sg._p(x), sg._p(w), b.data_ptr(), None, sg._p(y),
                                        sg._p(hw), sg._p(hb), sg._p(hy), 1, 16, 16, 32, 8, 1,
                                        sg._stream()) == 0


def test_unet_inference_forward_equals_training_forward(sg):
    """The inference forward (decoder inputs read without concatenation)
    returns the training forward's bits."""
    dev = "cuda:0"
    net = sg.UNet(seed=6, device=dev)
    g = torch.Generator().manual_seed(11)
    x = (torch.randn(2, 64, 48, 8, generator=g) * 0.5).to(dev).to(torch.bfloat16).contiguous()
    y1 = net.forward(x, keep=True)
    y2 = net.forward(x, keep=False)
    assert torch.equal(y1.view(torch.int32), y2.view(torch.int32))
    with pytest.raises(RuntimeError):
        net.backward(torch.zeros(2, 64, 48, 8, dtype=torch.bfloat16, device=dev))


def test_unet_training_reduces_loss(sg):
    """A few Adam steps on one fixed batch must cut the loss (end-to-end wiring)."""
    dev = "cuda:0"
    net = sg.UNet(seed=1, device=dev)
    g = torch.Generator().manual_seed(2)
    x = (torch.randn(4, 32, 32, 8, generator=g) * 0.5).to(dev).to(torch.bfloat16)
    t = torch.tanh(x[..., :3].float() * 2.0)
    losses = []
    for _ in range(30):
        y = net.forward(x)
        loss, gy = sg.mse_loss_and_grad(y, t)
        net.backward(gy)
        net.step(3e-3)
        losses.append(loss.item())
    assert losses[-1] < 0.3 * losses[0], losses
