"""Learned gradient preconditioner: a 4-level U-Net on the HIP kernels of
csrc/unet.hip (BF16 weights and activations, FP32 accumulation on MFMA).

It maps the raw FP32 gradient and the current model of an FWI iteration to an
update direction.  There is no counterpart in the reference: its FWI step
(seistorch/fwi.py with the optimizers of seistorch/optimizer.py) applies the
plain gradient, which remains the fallback whenever the network's direction
does not reduce the misfit.

Architecture (NHWC; C = 32):
    input 2 channels (gradient * gscale, (model - mshift) * mscale), padded to 8
    encoder level l = 0..3 at 1/2^l resolution: conv3x3 -> GELU -> conv3x3 -> GELU
        with C, 2C, 4C, 8C channels, 2x2 average pooling between levels
    bottleneck at 1/16: two conv3x3 + GELU with 8C channels
    decoder l = 3..0: nearest 2x upsampling, concatenation with the encoder
        output of level l, two conv3x3 + GELU back to the level's channels
    head: 1x1 convolution to one FP32 channel
H and W must be multiples of 16.
"""
import math

import numpy as np
import torch

from . import _lib

BASE = 32
IN_CH = 8  # 2 used, padded for 16-byte vector access


def layer_specs(base=BASE):
    """[(name, Cin, Cout)] of the 3x3 convolutions in forward order."""
    ch = [base, 2 * base, 4 * base, 8 * base]
    specs, cin = [], IN_CH
    for l, c in enumerate(ch):
        specs += [(f"enc{l}a", cin, c), (f"enc{l}b", c, c)]
        cin = c
    specs += [("bota", cin, ch[3]), ("botb", ch[3], ch[3])]
    low = ch[3]
    for l in (3, 2, 1, 0):
        specs += [(f"dec{l}a", low + ch[l], ch[l]), (f"dec{l}b", ch[l], ch[l])]
        low = ch[l]
    return specs


class Params:
    """FP32 master parameters in one flat arena (conv weights (Cout, 9*Cin)
    then bias (Cout) per layer, then head weight (C) and bias (1)), plus the
    bf16 copy of the conv weights the kernels read."""

    def __init__(self, device, base=BASE, seed=0):
        self.specs = layer_specs(base)
        self.base = base
        self.off = {}
        o = 0
        for name, cin, cout in self.specs:
            self.off[name] = (o, o + cout * 9 * cin)  # weight, bias offsets
            o += cout * 9 * cin + cout
        self.off["head"] = (o, o + base)
        o += base + 1
        self.n = o
        g = torch.Generator().manual_seed(seed)
        host = torch.zeros(self.n)
        for name, cin, cout in self.specs:
            w0, _ = self.off[name]
            fan_in = 9 * (2 if name == "enc0a" else cin)
            w = torch.randn(cout, 9, cin, generator=g) * math.sqrt(2.0 / fan_in)
            if name == "enc0a":
                w[:, :, 2:] = 0.0  # the padding channels carry nothing
            host[w0:w0 + cout * 9 * cin] = w.reshape(-1)
        h0, _ = self.off["head"]
        host[h0:h0 + base] = torch.randn(base, generator=g) * (0.1 / math.sqrt(base))
        self.master = host.to(device)
        self.bf16 = torch.empty(self.n, dtype=torch.int16, device=device)
        self.refresh_bf16()

    def refresh_bf16(self):
        _lib.call("seis_unet_to_bf16", _lib.ptr(self.master), self.n, _lib.ptr(self.bf16),
                  _lib.stream_of(self.master))

    def weight(self, name, arena=None):
        a = self.master if arena is None else arena
        w0, b0 = self.off[name]
        return a[w0:b0]

    def bias(self, name, arena=None):
        a = self.master if arena is None else arena
        w0, b0 = self.off[name]
        cout = dict((s[0], s[2]) for s in self.specs)[name]
        return a[b0:b0 + cout]

    def wbf16(self, name):
        w0, b0 = self.off[name]
        return self.bf16[w0:b0]

    def head(self, arena=None):
        a = self.master if arena is None else arena
        h0, hb = self.off["head"]
        return a[h0:hb], a[hb:hb + 1]

    def state_dict(self):
        return {"master": self.master.detach().cpu(), "base": self.base}

    def load_state_dict(self, sd):
        self.master.copy_(sd["master"].to(self.master.device))
        self.refresh_bf16()


def _bf(shape, dev):
    return torch.empty(shape, dtype=torch.int16, device=dev)


def _f32(shape, dev):
    return torch.empty(shape, dtype=torch.float32, device=dev)


class UNet:
    """Forward / backward of the network on (N, H, W) gradient and model planes."""

    def __init__(self, params):
        self.p = params

    # -- building blocks -------------------------------------------------
    def _conv(self, x, N, H, W, name, keep):
        _, cin, cout = next(s for s in self.p.specs if s[0] == name)
        dev = x.device
        y = _bf((N, H, W, cout), dev)
        z = _bf((N, H, W, cout), dev) if keep else None
        _lib.call("seis_unet_conv3x3", _lib.ptr(x), N, H, W, cin, _lib.ptr(self.p.wbf16(name)),
                  _lib.ptr(self.p.bias(name)), cout, 1, _lib.ptr(z), _lib.ptr(y), None,
                  _lib.stream_of(x))
        if keep:
            self.saved[name] = (x, z, (N, H, W))
        return y

    def forward(self, grad, model, scales, keep=False):
        """grad, model (N, H, W) FP32; scales (N, 3) FP32 = (gscale, mscale, mshift).
        Returns the (N, H, W) FP32 direction."""
        N, H, W = grad.shape
        if H % 16 or W % 16:
            raise RuntimeError(f"U-Net input {H}x{W} must be a multiple of 16 in both axes")
        _lib.require_device(grad, model, scales)
        dev = grad.device
        s = _lib.stream_of(grad)
        self.saved = {}
        x = _bf((N, H, W, IN_CH), dev)
        _lib.call("seis_unet_pack_input", _lib.ptr(grad), _lib.ptr(model), N, H, W,
                  _lib.ptr(scales), _lib.ptr(x), s)
        skips, h, w = [], H, W
        for l in range(4):
            x = self._conv(x, N, h, w, f"enc{l}a", keep)
            x = self._conv(x, N, h, w, f"enc{l}b", keep)
            skips.append((x, h, w))
            c = x.shape[-1]
            p = _bf((N, h // 2, w // 2, c), dev)
            _lib.call("seis_unet_avgpool2", _lib.ptr(x), N, h, w, c, _lib.ptr(p), s)
            x, h, w = p, h // 2, w // 2
        x = self._conv(x, N, h, w, "bota", keep)
        x = self._conv(x, N, h, w, "botb", keep)
        for l in (3, 2, 1, 0):
            sk, h2, w2 = skips[l]
            c1, c2 = x.shape[-1], sk.shape[-1]
            u = _bf((N, h2, w2, c1 + c2), dev)
            _lib.call("seis_unet_up_concat", _lib.ptr(x), _lib.ptr(sk), N, h2, w2, c1, c2,
                      _lib.ptr(u), s)
            h, w = h2, w2
            x = self._conv(u, N, h, w, f"dec{l}a", keep)
            x = self._conv(x, N, h, w, f"dec{l}b", keep)
        out = _f32((N, H, W), dev)
        hw, hb = self.p.head()
        _lib.call("seis_unet_head", _lib.ptr(x), N * H * W, self.p.base, _lib.ptr(hw),
                  _lib.ptr(hb), _lib.ptr(out), s)
        if keep:
            self.saved["head_in"] = x
            self.saved["shape"] = (N, H, W)
            self.saved["skip_c"] = [sk.shape[-1] for sk, _, _ in skips]
        return out

    def _conv_bwd(self, name, dy, gw_arena, need_dx):
        """dy (N,H,W,Cout) FP32 of the GELU output -> accumulates dW, db;
        returns dx (N,H,W,Cin) FP32 if need_dx."""
        _, cin, cout = next(s for s in self.p.specs if s[0] == name)
        x, z, (N, H, W) = self.saved[name]
        dev = dy.device
        s = _lib.stream_of(dy)
        dz = _bf((N, H, W, cout), dev)
        _lib.call("seis_unet_gelu_bwd", _lib.ptr(dy), _lib.ptr(z), dy.numel(), _lib.ptr(dz), s)
        _lib.call("seis_unet_conv3x3_wgrad", _lib.ptr(x), _lib.ptr(dz), N, H, W, cin, cout,
                  _lib.ptr(self.p.weight(name, gw_arena)), _lib.ptr(self.p.bias(name, gw_arena)),
                  s)
        if not need_dx:
            return None
        wt = _bf((cin, 9 * cout), dev)
        _lib.call("seis_unet_flip_weights", _lib.ptr(self.p.wbf16(name)), cout, cin,
                  _lib.ptr(wt), s)
        dx = _f32((N, H, W, cin), dev)
        _lib.call("seis_unet_conv3x3", _lib.ptr(dz), N, H, W, cout, _lib.ptr(wt), None, cin, 0,
                  None, None, _lib.ptr(dx), s)
        return dx

    def backward(self, dout, grad_arena):
        """dout (N, H, W) FP32 -> parameter gradients added into grad_arena
        (same layout as the master arena; the caller zeroes it)."""
        N, H, W = self.saved["shape"]
        dev = dout.device
        s = _lib.stream_of(dout)
        C = self.p.base
        xh = self.saved["head_in"]
        dx = _f32((N, H, W, C), dev)
        hw, _ = self.p.head()
        gw, gb = self.p.head(grad_arena)
        _lib.call("seis_unet_head_bwd", _lib.ptr(xh), _lib.ptr(dout.contiguous()), N * H * W, C,
                  _lib.ptr(hw), _lib.ptr(dx), _lib.ptr(gw), _lib.ptr(gb), s)
        dskip = [None] * 4
        dy = dx
        # decoder, finest level first (reverse of the forward order)
        for l in (0, 1, 2, 3):
            dy = self._conv_bwd(f"dec{l}b", dy, grad_arena, True)
            du = self._conv_bwd(f"dec{l}a", dy, grad_arena, True)
            _, _, (n_, h2, w2) = self.saved[f"dec{l}a"]
            c2 = self.saved["skip_c"][l]
            c1 = du.shape[-1] - c2
            dlow = _f32((N, h2 // 2, w2 // 2, c1), dev)
            dskip[l] = _f32((N, h2, w2, c2), dev)
            _lib.call("seis_unet_up_concat_bwd", _lib.ptr(du), N, h2, w2, c1, c2,
                      _lib.ptr(dlow), _lib.ptr(dskip[l]), 0, s)
            dy = dlow
        dy = self._conv_bwd("botb", dy, grad_arena, True)
        dy = self._conv_bwd("bota", dy, grad_arena, True)
        for l in (3, 2, 1, 0):
            # gradient of the encoder output: skip path + pooled path
            _, _, (n_, h, w) = self.saved[f"enc{l}b"]
            c = dskip[l].shape[-1]
            _lib.call("seis_unet_avgpool2_bwd", _lib.ptr(dy), N, h, w, c, _lib.ptr(dskip[l]), 1,
                      s)
            dy = self._conv_bwd(f"enc{l}b", dskip[l], grad_arena, True)
            dy = self._conv_bwd(f"enc{l}a", dy, grad_arena, l > 0)
        self.saved = {}


class AdamW:
    """torch.optim.AdamW semantics (decoupled weight decay) on the flat arena."""

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=1e-2):
        self.p = params
        self.lr, self.betas, self.eps, self.wd = lr, betas, eps, weight_decay
        self.m = torch.zeros_like(params.master)
        self.v = torch.zeros_like(params.master)
        self.t = 0

    def step(self, grad):
        self.t += 1
        # the bf16 arena is refreshed in the same pass
        _lib.call("seis_unet_adamw", _lib.ptr(self.p.master), _lib.ptr(grad), _lib.ptr(self.m),
                  _lib.ptr(self.v), self.p.n, self.lr, self.betas[0], self.betas[1], self.eps,
                  self.wd, self.t, _lib.ptr(self.p.bf16), _lib.stream_of(grad))


def mse(pred, target, mask=None):
    """(loss tensor (1,), dpred) of mean((mask*(pred-target))^2)."""
    loss = _f32((1,), pred.device)
    dpred = torch.empty_like(pred)
    n = pred.numel()
    count = float(mask.sum()) if mask is not None else float(n)
    _lib.call("seis_unet_mse", _lib.ptr(pred), _lib.ptr(target), _lib.ptr(mask), n,
              1.0 / max(count, 1.0), _lib.ptr(loss), _lib.ptr(dpred), _lib.stream_of(pred))
    return loss, dpred


def train_step(net, opt, grad, model, scales, target, mask=None):
    """One AdamW step of the network on a batch; returns the loss tensor."""
    pred = net.forward(grad, model, scales, keep=True)
    loss, dpred = mse(pred, target, mask)
    g = torch.empty_like(net.p.master)
    _lib.call("seis_unet_fill", _lib.ptr(g), g.numel(), 0.0, _lib.stream_of(g))
    net.backward(dpred, g)
    opt.step(g)
    return loss


def input_scales(grad, model, vref=3000.0, vspan=1500.0):
    """Per-image (gscale, mscale, mshift): gradient normalised by its max |g|,
    model mapped to (v - vref) / vspan."""
    gmax = grad.abs().flatten(1).amax(1).clamp_min(1e-30)
    N = grad.shape[0]
    sc = torch.empty((N, 3), device=grad.device, dtype=torch.float32)
    sc[:, 0] = 1.0 / gmax
    sc[:, 1] = 1.0 / vspan
    sc[:, 2] = vref
    return sc
