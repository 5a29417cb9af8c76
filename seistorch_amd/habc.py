"""HABC strip masks and blend weights (host side, computed once per geometry).

Same construction and FP32 values as seistorch/habc.py:4-81: 50-cell strips
with trapezoid masks (row r of the top strip keeps columns r..nx-1-r), weights
linspace(0, N, N)/N flipped (outer edge 1, row 49 0), assigned top, bottom,
left, right in that order; with a free surface (multiple) there is no top
strip and the side strips cover the top 50 rows fully.
"""
import torch


def bound_mask(nz, nx, w, dev="cpu", batchsize=1, return_idx=False, multiple=False):
    i = torch.arange(w, device=dev)[:, None]
    j = torch.arange(nx, device=dev)[None, :]
    top = ((j >= i) & (j <= nx - 1 - i)).float()
    bottom = torch.flip(top, [0])
    r = torch.arange(nz, device=dev)[:, None]
    k = torch.arange(w, device=dev)[None, :]
    left = ((k <= r) & (k <= nz - 1 - r)).float()
    right = torch.flip(left, [1])
    if multiple:
        top = None
        left[:w] = 1.0
        right[:w] = 1.0
    if not return_idx:
        return top, bottom, left, right
    rep = (lambda m: m.repeat(batchsize, 1, 1) == 1) if batchsize > 1 else (lambda m: m == 1)
    return (None if top is None else rep(top)), rep(bottom), rep(left), rep(right)


def generate_habc_coefficients_2d(domain_shape, N=50, multiple=False, device="cpu"):
    nz, nx = domain_shape
    d = torch.zeros(nz, nx, device=device)
    vals = torch.flip(torch.linspace(0.0, N, N, device=device) / N, [0])
    top, bottom, left, right = bound_mask(nz, nx, N, device, multiple=multiple)
    if N > 0:
        if not multiple:
            m = top == 1
            d[:N][m] = (vals[:, None].expand(N, nx) * top)[m]
        m = bottom == 1
        d[-N:][m] = (torch.flip(vals, [0])[:, None].expand(N, nx) * bottom)[m]
        m = left == 1
        d[:, :N][m] = (vals[None, :].expand(nz, N) * left)[m]
        m = right == 1
        d[:, -N:][m] = (torch.flip(vals, [0])[None, :].expand(nz, N) * right)[m]
    return d
