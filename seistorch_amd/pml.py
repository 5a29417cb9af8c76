"""PML damping coefficients (host side, computed once per geometry).

Same profile and FP32 arithmetic as seistorch/pml.py:29-59: the quadratic
profile d0*linspace(0, 1, N+1)**2 flipped, d0 = (1.5*1500/N)*log10(1e4), laid
along both ends of x and z (the top z strip omitted with a free surface) and
combined as sqrt(dx^2 + dz^2).
"""
import numpy as np
import torch


def generate_pml_coefficients_2d(domain_shape, N=50, B=100.0, multiple=False):
    nz, nx = domain_shape
    R = 1e-4
    order = 2
    cp = 1500.0
    d0 = (1.5 * cp / N) * np.log10(R ** -1)
    d_vals = torch.flip(d0 * torch.linspace(0.0, 1.0, N + 1) ** order, [0])
    pz = torch.zeros(nz)
    px = torch.zeros(nx)
    if N > 0:
        if not multiple:
            pz[0:N + 1] = d_vals
        pz[nz - N - 1:nz] = torch.flip(d_vals, [0])
        px[0:N + 1] = d_vals
        px[nx - N - 1:nx] = torch.flip(d_vals, [0])
    return torch.sqrt(px[None, :] ** 2 + pz[:, None] ** 2).to(torch.float32)
