"""CPU checks of the U-Net oracle: its weight/tap convention against a
direct loop, its bf16 rounding against known values, and its autograd
gradient against central finite differences (FP64, rounding off)."""
import numpy as np
import pytest
import torch

from oracle import unet as ou
from seistorch_amd.precond import layer_specs


def _offsets(specs, base):
    off, o = {}, 0
    for name, cin, cout in specs:
        off[name] = (o, o + cout * 9 * cin)
        o += cout * 9 * cin + cout
    off["head"] = (o, o + base)
    return off, o + base + 1


def test_conv_weight_layout_matches_direct_loop():
    rng = np.random.default_rng(0)
    N, H, W, cin, cout = 2, 5, 6, 8, 3
    x = rng.standard_normal((N, cin, H, W))
    arena = rng.standard_normal(cout * 9 * cin + cout)
    w, b = ou.conv_weight(torch.from_numpy(arena), (0, cout * 9 * cin), cin, cout, False)
    got = torch.nn.functional.conv2d(torch.from_numpy(x), w, b, padding=1).numpy()
    wl = arena[:cout * 9 * cin].reshape(cout, 9, cin)
    ref = np.zeros((N, cout, H, W))
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        patch = xp[:, :, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
        ref += np.einsum("nchw,oc->nohw", patch, wl[:, tap, :])
    ref += arena[cout * 9 * cin:][None, :, None, None]
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_bf16_rounding_known_values():
    one = 1.0
    vals = torch.tensor([one + 2 ** -8, one + 3 * 2 ** -8, one + 2 ** -9, -2.0 - 2 ** -7 - 2 ** -9,
                         3.0e38], dtype=torch.float32)
    want = torch.tensor([one, one + 2 ** -6, one, -2.0 - 2 ** -6, 2.9976e38])
    got = ou.bf16r(vals)
    assert torch.equal(got[:4], want[:4])  # ties to even; below-half rounds down
    assert abs(float(got[4]) / 3.0e38 - 1) < 2 ** -8


def test_oracle_gradient_matches_finite_differences():
    base = 8
    specs = layer_specs(base)
    off, n = _offsets(specs, base)
    g = torch.Generator().manual_seed(1)
    master = (torch.randn(n, generator=g, dtype=torch.float64) * 0.2).requires_grad_(True)
    N, H, W = 1, 16, 16
    grad = torch.randn((N, H, W), generator=g, dtype=torch.float64)
    model = 3000 + 300 * torch.randn((N, H, W), generator=g, dtype=torch.float64)
    scales = torch.tensor([[1.0, 1 / 1500.0, 3000.0]], dtype=torch.float64)
    tgt = torch.randn((N, H, W), generator=g, dtype=torch.float64)

    def loss(m):
        out = ou.forward(m, specs, off, base, grad, model, scales, emulate=False)
        return ((out - tgt) ** 2).mean()

    L = loss(master)
    L.backward()
    idx = [off["enc0a"][0] + 5, off["enc2b"][0] + 17, off["botb"][1] + 3, off["dec1a"][0] + 40,
           off["dec0b"][1], off["head"][0] + 2, n - 1]
    eps = 1e-6
    for i in idx:
        mp = master.detach().clone(); mp[i] += eps
        mm = master.detach().clone(); mm[i] -= eps
        fd = (loss(mp) - loss(mm)).item() / (2 * eps)
        assert abs(fd - master.grad[i].item()) <= 1e-6 + 1e-5 * abs(fd), (i, fd,
                                                                        master.grad[i].item())
