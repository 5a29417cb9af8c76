"""FP32 CPU restatement of the learned preconditioner (the U-Net of
seistorch_amd/precond.py), with the bf16 rounding of the HIP path emulated
at the same points (weights, network input, every stored activation).

TEST INFRASTRUCTURE: imported only by tests/, smoke() and bench.py; never by
the product.  The network has no counterpart in the reference, so this module
is its oracle: torch.nn.functional on the CPU, with autograd for gradients.
Rounding is straight-through (identity in the backward), so gradients are
those of the FP32 network evaluated at the rounded values.
"""
import torch
import torch.nn.functional as F


def bf16r(t, on=True):
    """Round to bf16 (nearest even) in the forward, identity in the backward."""
    if not on:
        return t
    return t + (t.to(torch.bfloat16).to(t.dtype) - t).detach()


def conv_weight(master, off, cin, cout, emulate=True):
    """(Cout, 9*Cin) arena slice -> (Cout, Cin, 3, 3) conv2d weight;
    tap = 3*(dy+1) + (dx+1)."""
    w0, b0 = off
    w = master[w0:b0].reshape(cout, 3, 3, cin).permute(0, 3, 1, 2)
    return bf16r(w, emulate), master[b0:b0 + cout]


def forward(master, specs, offsets, base, grad, model, scales, emulate=True, keep=None):
    """grad, model (N, H, W); scales (N, 3) = (gscale, mscale, mshift).
    Returns the (N, H, W) output (and fills `keep` with activations if a dict)."""
    conv = {name: (cin, cout) for name, cin, cout in specs}

    def c3(x, name):
        cin, cout = conv[name]
        w, b = conv_weight(master, offsets[name], cin, cout, emulate)
        z = F.conv2d(x, w, b, padding=1)
        y = bf16r(F.gelu(z), emulate)
        if keep is not None:
            keep[name] = (x, z, y)
        return y

    N, H, W = grad.shape
    g = grad * scales[:, 0, None, None]
    m = (model - scales[:, 2, None, None]) * scales[:, 1, None, None]
    x = torch.zeros((N, 8, H, W), dtype=grad.dtype)
    x = torch.cat([bf16r(g, emulate)[:, None], bf16r(m, emulate)[:, None],
                   torch.zeros((N, 6, H, W), dtype=grad.dtype)], 1)
    skips = []
    for l in range(4):
        x = c3(c3(x, f"enc{l}a"), f"enc{l}b")
        skips.append(x)
        x = bf16r(F.avg_pool2d(x, 2), emulate)
    x = c3(c3(x, "bota"), "botb")
    for l in (3, 2, 1, 0):
        u = torch.cat([F.interpolate(x, scale_factor=2, mode="nearest"), skips[l]], 1)
        x = c3(c3(u, f"dec{l}a"), f"dec{l}b")
    h0, hb = offsets["head"]
    return torch.einsum("nchw,c->nhw", x, master[h0:hb]) + master[hb]
