"""CPU pinning of the variable-density restatement (oracle/ref_torch.py:
rho_operator, rho_habc_step; acoustic_rho_habc.py:31-57)."""
import numpy as np
import pytest
import torch

from oracle import ref_torch
from seistorch_amd.habc import generate_habc_coefficients_2d


def _rng_fields(nz, nx, seed):
    g = torch.Generator().manual_seed(seed)
    rho = 1000.0 + 1500.0 * torch.rand((nz, nx), generator=g, dtype=torch.float64)
    p = torch.randn((2, nz, nx), generator=g, dtype=torch.float64)
    q = torch.randn((2, nz, nx), generator=g, dtype=torch.float64)
    return rho, p, q


def test_operator_is_symmetric():
    """div(b grad) with face-averaged buoyancy and zero padding is a
    symmetric matrix: <A p, q> = <p, A q>.  A swapped face weight, a sign or
    an index shift breaks this (planted below)."""
    rho, p, q = _rng_fields(13, 17, 0)
    A = ref_torch.rho_operator
    lhs, rhs = (A(rho, p) * q).sum(), (p * A(rho, q)).sum()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # planted: the x faces' weights swapped
    def wrong(rho, p):
        bu = rho.pow(-1)
        sh = ref_torch._shift
        return ((bu + sh(bu, 0, -1)) / 2 * (sh(p, 0, 1) - p) - (sh(bu, 0, 1) + bu) / 2 * (p - sh(p, 0, -1))
                + (bu + sh(bu, 1, 0)) / 2 * (sh(p, 1, 0) - p) - (sh(bu, -1, 0) + bu) / 2 * (p - sh(p, -1, 0)))
    assert abs((wrong(rho, p) * q).sum() - (p * wrong(rho, q)).sum()) > 1e-6 * abs(lhs)


def test_operator_is_negative_semidefinite_and_kills_constants_inside():
    rho, p, _ = _rng_fields(11, 12, 1)
    assert (ref_torch.rho_operator(rho, p) * p).sum() < 0
    one = torch.ones((1, 11, 12), dtype=torch.float64)
    inner = ref_torch.rho_operator(rho, one)[0, 1:-1, 1:-1]
    assert inner.abs().max() == 0.0


@pytest.mark.parametrize("multiple", [False, True])
def test_constant_density_reduces_to_the_acoustic_habc_step(multiple):
    """rho constant: b_pos = b_neg = 1/rho inside and rho * (1/rho) = 1, so
    the update is the constant-density one; on the outermost cells the
    buoyancy's zero padding differs, but every such cell lies in a strip of
    weight 1 (nothing of y survives) except the free-surface row."""
    nz, nx = 110, 120
    g = torch.Generator().manual_seed(2)
    c = 1500.0 + 1500.0 * torch.rand((nz, nx), generator=g, dtype=torch.float64)
    rho = torch.full((nz, nx), 2000.0, dtype=torch.float64)
    h1 = torch.randn((2, nz, nx), generator=g, dtype=torch.float64)
    h2 = torch.randn((2, nz, nx), generator=g, dtype=torch.float64)
    b = torch.as_tensor(generate_habc_coefficients_2d((nz, nx), 50, multiple=multiple),
                        dtype=torch.float64)
    dt, h = 1e-3, 10.0
    y_rho, _ = ref_torch.rho_habc_step(c, rho, h1, h2, dt, h, b, multiple=multiple)
    y_ac, _ = ref_torch.habc_step(c, h1, h2, torch.tensor(dt, dtype=torch.float64),
                                  torch.tensor(h, dtype=torch.float64), b, multiple=multiple)
    rows = slice(1, None) if multiple else slice(None)
    diff = (y_rho - y_ac)[:, rows].abs().max() / y_ac.abs().max()
    assert diff < 1e-13
    if multiple:  # the free surface row keeps y: there the padding shows
        assert (y_rho - y_ac)[:, 0].abs().max() > 0
