"""Convolutional surrogate that warm-starts discharge and momentum.

No reference counterpart: the reference converges the exponentially averaged
discharge/momentum fields (source/world.h:80-86) only by running the particle
method.  This module predicts the converged fields from a terrain so the
particle method can start near its fixed point.  The particle method stays
the fallback and the arbiter (``train_surrogate.warm_start``).

The network is a 4-level residual U-Net, 3x3 convolutions,
32-64-96-128 channels, BF16 weights and activations with FP32 accumulation.
Every operation -- forward, backward (data and weight gradients), pooling,
upsampling, loss and Adam -- is a HIP kernel of ``libshy.so``
(``csrc/shy_nn.hip``, BF16 MFMA for all convolutions).  torch supplies device
memory and the stream only.

Input channels (NHWC, 8): height centred on the window mean (x4), the
central-difference slopes along x and y (height per cell x 512), the
transformed discharge and momentum of :func:`encode_flow` of the current
state, two zero channels.  Output: converged (discharge, momentumx,
momentumy) in the transformed space (the head has 8 outputs; the last 5 are
unused).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import lib as _shy_lib

_nn = None


def nn():
    global _nn
    if _nn is None:
        L = _shy_lib()
        vp, i32, f32, sz = C.c_void_p, C.c_int, C.c_float, C.c_size_t
        sig = {
            "shy_nn_last_error": (C.c_char_p, []),
            "shy_nn_conv_fwd": (i32, [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, i32,
                                      vp]),
            "shy_nn_conv_fwd_upcat": (i32, [vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32,
                                            i32, i32, vp]),
            "shy_nn_conv_dgrad": (i32, [vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, vp]),
            "shy_nn_conv_fwd_head": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, i32,
                                           i32, i32, vp]),
            "shy_nn_conv_fwd_pool": (i32, [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, i32,
                                           vp]),
            "shy_nn_flip_weights": (i32, [vp, vp, i32, i32, i32, vp]),
            "shy_nn_conv_wgrad": (i32, [vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, i32, vp]),
            "shy_nn_conv_wgrad_noalloc": (i32, [vp, vp, vp, i32, i32, i32, i32, i32, i32, vp]),
            "shy_nn_bias_grad": (i32, [vp, vp, vp, i32, i32, i32, vp]),
            "shy_nn_relu_bwd": (i32, [vp, vp, vp, sz, vp]),
            "shy_nn_pool2": (i32, [vp, vp, i32, i32, i32, i32, i32, vp]),
            "shy_nn_up2": (i32, [vp, vp, i32, i32, i32, i32, i32, i32, i32, vp]),
            "shy_nn_channels": (i32, [vp, vp, sz, i32, i32, i32, i32, vp]),
            "shy_nn_add": (i32, [vp, vp, vp, sz, vp]),
            "shy_nn_mse": (i32, [vp, vp, vp, vp, vp, sz, f32, vp]),
            "shy_nn_adam": (i32, [vp, vp, vp, vp, vp, sz, f32, f32, f32, f32, i32, f32, vp]),
            "shy_nn_adam_conv": (i32, [vp, vp, vp, vp, vp, vp, i32, i32, i32, f32, f32, f32, f32,
                                       i32, f32, vp]),
            "shy_nn_cast": (i32, [vp, vp, sz, i32, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _nn = L
    return _nn


def _ck(rc):
    if rc != 0:
        raise RuntimeError("shy_nn: " + nn().shy_nn_last_error().decode())


def _p(t):
    return None if t is None else t.data_ptr()


def _stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


# ------------------------------------------------------------------ encoding
D_SCALE = 5.0


def encode_flow(d, mx, my):
    """(discharge, momentum) -> network space: log1p(d)/5, signed log1p(|m|)/5."""
    import torch
    return torch.stack([torch.log1p(d.clamp_min(0)) / D_SCALE,
                        torch.sign(mx) * torch.log1p(mx.abs()) / D_SCALE,
                        torch.sign(my) * torch.log1p(my.abs()) / D_SCALE], -1)


def decode_flow(y):
    import torch
    d = torch.expm1((y[..., 0] * D_SCALE).clamp(max=20.0)).clamp_min(0)
    mx = torch.sign(y[..., 1]) * torch.expm1((y[..., 1].abs() * D_SCALE).clamp(max=20.0))
    my = torch.sign(y[..., 2]) * torch.expm1((y[..., 2].abs() * D_SCALE).clamp(max=20.0))
    return d, mx, my


SLOPE_SCALE = 512.0
H_SCALE = 4.0


def features(h, d, mx, my):
    """Per-cell input features of a whole map ([X, Y] f32 device tensors) ->
    [X, Y, 8] f32.  Channel 0 holds the raw height; :func:`window_batch`
    centres it per window."""
    import torch
    X, Y = h.shape
    f = torch.zeros(X, Y, 8, dtype=torch.float32, device=h.device)
    f[..., 0] = h
    hp = torch.nn.functional.pad(h[None, None], (1, 1, 1, 1), mode="replicate")[0, 0]
    f[..., 1] = (hp[2:, 1:-1] - hp[:-2, 1:-1]) * SLOPE_SCALE
    f[..., 2] = (hp[1:-1, 2:] - hp[1:-1, :-2]) * SLOPE_SCALE
    f[..., 3:6] = encode_flow(d, mx, my)
    return f


def window_batch(feats, origins, tile):
    """Stack windows feats[x0:x0+tile, y0:y0+tile] into a bf16 NHWC batch,
    with the height channel centred and scaled per window."""
    import torch
    xs = torch.stack([feats[x0:x0 + tile, y0:y0 + tile] for x0, y0 in origins])
    h = xs[..., 0]
    xs[..., 0] = (h - h.mean(dim=(1, 2), keepdim=True)) * H_SCALE
    return xs.to(torch.bfloat16).contiguous()


def make_input(h, d, mx, my):
    """One window covering the whole map: [1, X, Y, 8] bf16."""
    return window_batch(features(h, d, mx, my), [(0, 0)], h.shape[0])[0]


# -------------------------------------------------------------------- layers
class Conv:
    """taps x taps convolution with FP32 master weights and a bf16 copy."""

    def __init__(self, cin, cout, taps, gen, device):
        import torch
        self.cin, self.cout, self.taps = cin, cout, taps
        fan_in = cin * taps
        std = math.sqrt(2.0 / fan_in)
        w = torch.from_numpy(gen.standard_normal((cout, taps * cin)).astype(np.float32) * std)
        self.w = w.to(device)
        self.b = torch.zeros(cout, dtype=torch.float32, device=device)
        self.w_bf = self.w.to(torch.bfloat16)
        self.wf_bf = torch.empty(cin * taps * cout, dtype=torch.bfloat16, device=device)
        self._wf_of = None  # (w_bf's address, its torch version) that wf_bf was flipped from
        self.dw = torch.zeros_like(self.w)
        self.db = torch.zeros_like(self.b)
        self.m = [torch.zeros_like(self.w), torch.zeros_like(self.b)]
        self.v = [torch.zeros_like(self.w), torch.zeros_like(self.b)]

    def forward(self, x, shape, res=None, relu=True, out_f32=False):
        import torch
        N, H, W = shape
        y = torch.empty(N, H, W, self.cout, device=x.device,
                        dtype=torch.float32 if out_f32 else torch.bfloat16)
        _ck(nn().shy_nn_conv_fwd(_p(x), _p(self.w_bf), _p(self.b), _p(res),
                                 None if out_f32 else _p(y), _p(y) if out_f32 else None,
                                 N, H, W, self.cin, self.cout, self.taps, 1 if relu else 0,
                                 _stream()))
        return y

    def forward_head(self, x, shape, head, res=None, relu=True, keep=True):
        """This 3x3 convolution (32 output channels) and the 1x1 convolution
        ``head`` (bias, no ReLU, FP32 output) applied to its rounded output,
        in one pass.  Returns (y or None without ``keep``, head output)."""
        import torch
        N, H, W = shape
        assert self.cout == 32 and self.taps == 9 and head.cin == 32 and head.taps == 1
        y = (torch.empty(N, H, W, self.cout, device=x.device, dtype=torch.bfloat16)
             if keep else None)
        out = torch.empty(N, H, W, head.cout, device=x.device, dtype=torch.float32)
        _ck(nn().shy_nn_conv_fwd_head(_p(x), _p(self.w_bf), _p(self.b), _p(res), _p(y),
                                      _p(head.w_bf), _p(head.b), _p(out), N, H, W, self.cin,
                                      head.cout, 1 if relu else 0, _stream()))
        return y, out

    def forward_pool(self, x, shape, res=None, relu=True):
        """This 3x3 convolution and the 2x2 average pool of its output in one
        pass (the bits of forward + shy_nn_pool2).  Returns (y, pooled y)."""
        import torch
        N, H, W = shape
        assert self.taps == 9
        y = torch.empty(N, H, W, self.cout, device=x.device, dtype=torch.bfloat16)
        p = torch.empty(N, H // 2, W // 2, self.cout, device=x.device, dtype=torch.bfloat16)
        _ck(nn().shy_nn_conv_fwd_pool(_p(x), _p(self.w_bf), _p(self.b), _p(res), _p(y), _p(p),
                                      N, H, W, self.cin, self.cout, 1 if relu else 0,
                                      _stream()))
        return y, p

    def forward_upcat(self, xh, skip, shape, res=None, relu=True):
        """The same convolution of [up2(xh), skip] along channels, read from
        the two tensors (no concatenated copy)."""
        import torch
        N, H, W = shape
        c1, c2 = xh.shape[-1], skip.shape[-1]
        assert c1 + c2 == self.cin and self.taps == 9
        y = torch.empty(N, H, W, self.cout, device=xh.device, dtype=torch.bfloat16)
        _ck(nn().shy_nn_conv_fwd_upcat(_p(xh), _p(skip), _p(self.w_bf), _p(self.b), _p(res),
                                       _p(y), None, N, H, W, c1, c2, self.cout,
                                       1 if relu else 0, _stream()))
        return y

    def backward(self, x, g, shape, scratch, need_dx=True, gres=None, mask=None):
        """g: gradient of the pre-activation output (bf16). Returns dx (bf16);
        with ``gres`` and ``mask`` (32-channel multiples) dx + gres, zeroed
        where mask <= 0, from the data-gradient kernel's epilogue."""
        import torch
        N, H, W = shape
        P = N * H * W
        part, splits = scratch.part_for(self, P)
        _ck(nn().shy_nn_conv_wgrad(_p(x), _p(g), _p(self.dw), _p(part), splits, N, H, W,
                                   self.cin, self.cout, self.taps, _stream()))
        _ck(nn().shy_nn_bias_grad(_p(g), _p(self.db), _p(scratch.bias_part), scratch.bias_rows,
                                  P, self.cout, _stream()))
        if not need_dx:
            return None
        if self._wf_of != (self.w_bf.data_ptr(), self.w_bf._version):
            # w_bf was replaced or written through torch since step() last
            # wrote wf_bf (or step() has not run): flip it here
            _ck(nn().shy_nn_flip_weights(_p(self.w_bf), _p(self.wf_bf), self.cin, self.cout,
                                         self.taps, _stream()))
            self._wf_of = (self.w_bf.data_ptr(), self.w_bf._version)
        dx = torch.empty(N, H, W, self.cin, dtype=torch.bfloat16, device=g.device)
        if gres is not None or mask is not None:
            _ck(nn().shy_nn_conv_dgrad(_p(g), _p(self.wf_bf), _p(gres), _p(mask), _p(dx), N, H, W,
                                       self.cin, self.cout, self.taps, _stream()))
            return dx
        _ck(nn().shy_nn_conv_fwd(_p(g), _p(self.wf_bf), None, None, _p(dx), None, N, H, W,
                                 self.cout, self.cin, self.taps, 0, _stream()))
        return dx

    def step(self, lr, t, wd=0.0):
        s = _stream()
        # the kernel writes w_bf and its flipped copy wf_bf (torch's version
        # counter of w_bf does not move: it counts writes made through torch)
        _ck(nn().shy_nn_adam_conv(_p(self.w), _p(self.dw), _p(self.m[0]), _p(self.v[0]),
                                  _p(self.w_bf), _p(self.wf_bf), self.cin, self.cout, self.taps,
                                  lr, 0.9, 0.999, 1e-8, t, wd, s))
        self._wf_of = (self.w_bf.data_ptr(), self.w_bf._version)
        _ck(nn().shy_nn_adam(_p(self.b), _p(self.db), _p(self.m[1]), _p(self.v[1]), None,
                             self.b.numel(), lr, 0.9, 0.999, 1e-8, t, 0.0, s))

    def params(self):
        return [self.w, self.b]


class Scratch:
    """FP32 partial-sum buffers for the weight and bias gradients."""

    def __init__(self, device, max_floats=1 << 24):
        import torch
        self.part = torch.empty(max_floats, dtype=torch.float32, device=device)
        self.bias_rows = 1024
        self.bias_part = torch.empty(self.bias_rows * 1024, dtype=torch.float32, device=device)

    def part_for(self, conv, P):
        tiles = ((conv.cout + 31) // 32) * ((conv.taps * conv.cin + 31) // 32)
        splits = max(1, min(P // 512, -(-4096 // tiles)))
        splits = min(splits, self.part.numel() // (conv.cout * conv.taps * conv.cin))
        return self.part, max(1, splits)


def _relu_bwd(y, g):
    import torch
    out = torch.empty_like(g)
    _ck(nn().shy_nn_relu_bwd(_p(y), _p(g), _p(out), g.numel(), _stream()))
    return out


def _add(a, b):
    import torch
    out = torch.empty_like(a)
    _ck(nn().shy_nn_add(_p(a), _p(b), _p(out), a.numel(), _stream()))
    return out


class ResStage:
    """conv_a (cin->c) ReLU, conv_b ReLU, conv_c + skip from a, ReLU."""

    def __init__(self, cin, c, gen, device):
        self.a = Conv(cin, c, 9, gen, device)
        self.b = Conv(c, c, 9, gen, device)
        self.c = Conv(c, c, 9, gen, device)

    def forward(self, x, shape, skip=None, head=None, pool=False):
        """With ``skip`` the stage's input is [up2(x), skip] along channels,
        read without building it (inference: nothing is kept for backward).
        With ``head`` (a 1x1 Conv) returns (output or None with ``skip``,
        head applied to it) from the last convolution's one pass.  With
        ``pool`` returns (output, its 2x2 average pool), also from one pass."""
        if skip is None:
            ya = self.a.forward(x, shape)
        else:
            ya = self.a.forward_upcat(x, skip, shape)
        yb = self.b.forward(ya, shape)
        pooled = None
        if pool:
            yc, pooled = self.c.forward_pool(yb, shape, res=ya)
        elif head is None:
            yc = self.c.forward(yb, shape, res=ya)
        else:
            yc, out = self.c.forward_head(yb, shape, head, res=ya, keep=skip is None)
        self.saved = (x, ya, yb, yc) if skip is None else None
        if pool:
            return yc, pooled
        return yc if head is None else (yc, out)

    def backward(self, g, shape, scratch, need_dx=True):
        x, ya, yb, yc = self.saved
        gc = _relu_bwd(yc, g)
        # relu_bwd(yb, dgrad_c) and relu_bwd(ya, gc + dgrad_b) in the
        # data-gradient epilogues (the same bits as the separate passes)
        gb = self.c.backward(yb, gc, shape, scratch, mask=yb)
        ga = self.b.backward(ya, gb, shape, scratch, gres=gc, mask=ya)
        return self.a.backward(x, ga, shape, scratch, need_dx=need_dx)

    def convs(self):
        return [self.a, self.b, self.c]


class UNet:
    CH = (32, 64, 96, 128)

    def __init__(self, seed=0, device="cuda:0"):
        gen = np.random.default_rng(seed)
        c0, c1, c2, c3 = self.CH
        self.e = [ResStage(8, c0, gen, device), ResStage(c0, c1, gen, device),
                  ResStage(c1, c2, gen, device), ResStage(c2, c3, gen, device)]
        self.d = [ResStage(c1 + c0, c0, gen, device), ResStage(c2 + c1, c1, gen, device),
                  ResStage(c3 + c2, c2, gen, device)]
        self.head = Conv(c0, 8, 1, gen, device)
        self.device = device
        self.scratch = Scratch(device)
        self.t = 0

    def convs(self):
        out = []
        for s in self.e + self.d:
            out += s.convs()
        return out + [self.head]

    def n_params(self):
        return sum(c.w.numel() + c.b.numel() for c in self.convs())

    def flops(self, N, H, W):
        """multiply-adds x 2 of one forward pass"""
        f = 0
        c0, c1, c2, c3 = self.CH
        sizes = [(H >> k) * (W >> k) for k in range(4)]
        for k, s in enumerate(self.e):
            for cv in s.convs():
                f += 2 * cv.taps * cv.cin * cv.cout * sizes[k]
        for k, s in enumerate(self.d):
            for cv in s.convs():
                f += 2 * cv.taps * cv.cin * cv.cout * sizes[k]
        f += 2 * self.head.cin * self.head.cout * sizes[0]
        return f * N

    def forward(self, x, keep=True):
        """x: [N, H, W, 8] bf16 (H, W multiples of 8). Returns [N, H, W, 8] f32.
        ``keep`` saves what :meth:`backward` needs; without it (inference) the
        decoder stages read their upsampled-and-concatenated input straight
        from the two tensors, with the same bits."""
        import torch
        N, H, W, _ = x.shape
        self.shapes = [(N, H >> k, W >> k) for k in range(4)]
        self.kept = keep
        skips = []
        h = x
        for k in range(4):
            if k < 3:
                # the stage's last convolution also writes the pooled input
                # of the next stage
                h, p = self.e[k].forward(h, self.shapes[k], pool=True)
                skips.append(h)
                h = p
            else:
                h = self.e[k].forward(h, self.shapes[k])
        self.cat = []
        for k in (2, 1, 0):
            n, hh, ww = self.shapes[k]
            cu, cs = h.shape[-1], skips[k].shape[-1]
            # the last stage's last convolution applies the head in its own pass
            head = self.head if k == 0 else None
            if not keep:
                h = self.d[k].forward(h, self.shapes[k], skip=skips[k], head=head)
                continue
            u = torch.empty(n, hh, ww, cu + cs, dtype=torch.bfloat16, device=x.device)
            _ck(nn().shy_nn_up2(_p(h), _p(u), n, hh // 2, ww // 2, cu, cu + cs, 0, 0, _stream()))
            _ck(nn().shy_nn_channels(_p(skips[k]), _p(u), n * hh * ww, cs, cu + cs, cu, 0,
                                     _stream()))
            self.cat.append((cu, cs))
            h = self.d[k].forward(u, self.shapes[k], head=head)
        self.h0, out = h
        return out

    def backward(self, gy):
        """gy: [N, H, W, 8] bf16 gradient of the output."""
        import torch
        if not self.kept:
            raise RuntimeError("backward needs a forward(..., keep=True)")
        g = self.head.backward(self.h0, gy, self.shapes[0], self.scratch)
        gskip = [None, None, None]
        for k, (cu, cs) in zip((0, 1, 2), reversed(self.cat)):
            n, hh, ww = self.shapes[k]
            gu = self.d[k].backward(g, self.shapes[k], self.scratch)
            gs = torch.empty(n, hh, ww, cs, dtype=torch.bfloat16, device=gy.device)
            _ck(nn().shy_nn_channels(_p(gu), _p(gs), n * hh * ww, cs, cu + cs, cu, 1, _stream()))
            gskip[k] = gs
            g = torch.empty(n, hh // 2, ww // 2, cu, dtype=torch.bfloat16, device=gy.device)
            _ck(nn().shy_nn_up2(_p(gu), _p(g), n, hh // 2, ww // 2, cu, cu + cs, 0, 1, _stream()))
        for k in (3, 2, 1, 0):
            gx = self.e[k].backward(g, self.shapes[k], self.scratch, need_dx=k > 0)
            if k == 0:
                break
            n, hh, ww = self.shapes[k - 1]
            c = gx.shape[-1]
            gp = torch.empty(n, hh, ww, c, dtype=torch.bfloat16, device=gy.device)
            _ck(nn().shy_nn_pool2(_p(gx), _p(gp), n, hh, ww, c, 1, _stream()))
            g = _add(gp, gskip[k - 1])

    def step(self, lr, wd=0.0):
        self.t += 1
        for cv in self.convs():
            cv.step(lr, self.t, wd)

    def state_dict(self):
        return {f"w{i}": c.w.detach().cpu().numpy() for i, c in enumerate(self.convs())} | \
               {f"b{i}": c.b.detach().cpu().numpy() for i, c in enumerate(self.convs())}

    def load_state_dict(self, sd):
        import torch
        for i, c in enumerate(self.convs()):
            c.w.copy_(torch.from_numpy(sd[f"w{i}"]))
            c.b.copy_(torch.from_numpy(sd[f"b{i}"]))
            c.w_bf.copy_(c.w.to(torch.bfloat16))


def mse_loss_and_grad(y, target, mask=None):
    """y [N,H,W,8] f32 head output; target [N,H,W,3] f32. Returns (loss, g bf16)."""
    import torch
    t8 = torch.zeros_like(y)
    t8[..., :3] = target
    m8 = torch.zeros_like(y)
    m8[..., :3] = 1.0 if mask is None else mask[..., None]
    count = float(m8.sum().item())
    g = torch.empty(y.shape, dtype=torch.bfloat16, device=y.device)
    nblk = -(-y.numel() // 256)
    part = torch.empty(nblk, dtype=torch.float32, device=y.device)
    _ck(nn().shy_nn_mse(_p(y), _p(t8), _p(m8), _p(g), _p(part), y.numel(), 1.0 / count,
                        _stream()))
    return part.sum() / count, g


def _tile_starts(n, tile, halo):
    """(window start, core start, core length) covering [0, n): cores of
    tile-2*halo cells, windows shifted inward at the map edges."""
    core = tile - 2 * halo
    out = []
    c0 = 0
    while c0 < n:
        clen = min(core, n - c0)
        w0 = min(max(c0 - halo, 0), n - tile)
        out.append((w0, c0, clen))
        c0 += clen
    return out


def predict_tiled(net, h, d, mx, my, tile=256, halo=32, batch=16):
    """Apply the network window-wise to a map of any size >= tile (inputs
    [X, Y] f32 device tensors): each window of tile x tile cells contributes
    its core, at least ``halo`` cells from every window edge that is not a map
    edge.  Returns decoded (d, mx, my)."""
    import torch
    X, Y = h.shape
    if X < tile or Y < tile:
        raise ValueError(f"map {X}x{Y} is smaller than one {tile}x{tile} window")
    feats = features(h, d, mx, my)
    out = torch.empty(X, Y, 3, dtype=torch.float32, device=h.device)
    jobs = [(wx, cx, lx, wy, cy, ly) for wx, cx, lx in _tile_starts(X, tile, halo)
            for wy, cy, ly in _tile_starts(Y, tile, halo)]
    for b0 in range(0, len(jobs), batch):
        js = jobs[b0:b0 + batch]
        y = net.forward(window_batch(feats, [(j[0], j[3]) for j in js], tile), keep=False)
        for k, (wx, cx, lx, wy, cy, ly) in enumerate(js):
            out[cx:cx + lx, cy:cy + ly] = y[k, cx - wx:cx - wx + lx, cy - wy:cy - wy + ly, :3]
    return decode_flow(out)
