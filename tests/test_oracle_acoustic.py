"""Pins the FP64 acoustic oracle with checks that do not restate its formulas.

* leapfrog standing wave: the zero-padded 5-point Laplacian has eigenmodes
  sin(pi m (i+1)/(n+1)); the undamped step must advance them by the exact
  discrete frequency 2cos(w dt) = 2 - c^2 dt^2 kappa.
* single damped cell (c = 0): the scheme is u'' + 2 d u' = 0 discretised, so a
  constant stays constant and the decaying root is (1 - d dt)/(1 + d dt).
* timing: a unit impulse at n = 0 is recorded at the source at r[0] = 1 and
  reaches a 5-point neighbour at r[1] = c^2 dt^2 / h^2 (rnn.py:178-202).
* reciprocity of the damped, heterogeneous scheme: G(r,s) c_s^2 = G(s,r) c_r^2.
* adjoint: central finite differences, the dot-product test, and torch
  autograd (FP64) of the op-for-op torch restatement.
"""
import numpy as np
import pytest
import torch

from oracle import acoustic as oa
from oracle import coeffs
from oracle import ref_torch


def _model(nz, nx, seed=0, vmin=1500.0, vmax=2500.0):
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, nz)[:, None]
    c = vmin + (vmax - vmin) * z + 0 * np.arange(nx)[None, :]
    c = c + 50.0 * rng.standard_normal((nz, nx))
    return c


def test_standing_wave_exact():
    nz, nx, h, dt, c0 = 20, 26, 10.0, 1e-3, 1500.0
    m, l = 3, 5
    zi = np.arange(nz)[:, None] + 1
    xi = np.arange(nx)[None, :] + 1
    phi = np.sin(np.pi * m * zi / (nz + 1)) * np.sin(np.pi * l * xi / (nx + 1))
    kappa = 4 / h ** 2 * (np.sin(np.pi * m / (2 * (nz + 1))) ** 2 + np.sin(np.pi * l / (2 * (nx + 1))) ** 2)
    cosw = 1 - 0.5 * c0 ** 2 * dt ** 2 * kappa
    omega = np.arccos(cosw)
    c = np.full((nz, nx), c0)
    d = np.zeros((nz, nx))
    u, up = phi[None].copy(), (np.cos(omega) * phi)[None]
    for n in range(1, 60):
        u, up = oa.step(c, u, up, d, dt, h), u
        np.testing.assert_allclose(u[0], np.cos(n * omega) * phi, atol=1e-10)


def test_damped_cell_roots():
    dt, dd = 1e-3, 137.0
    c = np.zeros((3, 3))
    d = np.full((3, 3), dd)
    one = np.ones((1, 3, 3))
    y = oa.step(c, one, one, d, dt, 10.0)
    np.testing.assert_allclose(y, one, rtol=1e-13)
    r = (1 - dd * dt) / (1 + dd * dt)
    u, up = r * one, one.copy()
    for n in range(2, 30):
        u, up = oa.step(c, u, up, d, dt, 10.0), u
        np.testing.assert_allclose(u, r ** n * one, rtol=1e-11)


def test_impulse_timing():
    nz, nx, h, dt, c0 = 12, 14, 10.0, 1e-3, 2000.0
    c = np.full((nz, nx), c0)
    d = np.zeros((nz, nx))
    w = np.zeros(6); w[0] = 1.0
    src = [[5, 6]]
    rec = [[[5, 6], [5, 7], [6, 6], [5, 9]]]
    r = oa.forward(c, d, dt, h, w, src, rec)[0][:, :, 0]
    assert r[0, 0] == 1.0 and np.all(r[0, 1:] == 0)
    np.testing.assert_allclose(r[1, 1], c0 ** 2 * dt ** 2 / h ** 2, rtol=1e-14)
    np.testing.assert_allclose(r[1, 2], c0 ** 2 * dt ** 2 / h ** 2, rtol=1e-14)
    assert np.all(r[:3, 3] == 0) and r[3, 3] != 0  # three cells away: first at n = 3


def test_zero_source_and_linearity():
    nz, nx = 30, 34
    c = _model(nz, nx, 1)
    d = coeffs.pml_coefficients_2d(nz, nx, N=8)
    w = coeffs.ricker(15.0, 1e-3, 80, delay=20)
    src = [[10, 12], [12, 20]]
    rec = [[[9, x] for x in range(nx)]] * 2
    r0 = oa.forward(c, d, 1e-3, 10.0, 0 * w, src, rec)
    assert all(np.all(x == 0) for x in r0)
    r1 = oa.forward(c, d, 1e-3, 10.0, w, src, rec)
    r2 = oa.forward(c, d, 1e-3, 10.0, -2.5 * w, src, rec)
    for a, b in zip(r1, r2):
        np.testing.assert_allclose(b, -2.5 * a, rtol=1e-12, atol=1e-12 * np.abs(a).max())


def test_reciprocity_damped_heterogeneous():
    nz, nx, h, dt = 32, 36, 10.0, 1e-3
    c = _model(nz, nx, 2)
    d = coeffs.pml_coefficients_2d(nz, nx, N=8)
    w = coeffs.ricker(15.0, dt, 120, delay=25)
    s, r = (11, 9), (17, 27)
    a = oa.forward(c, d, dt, h, w, [s], [[r]])[0][:, 0, 0]
    b = oa.forward(c, d, dt, h, w, [r], [[s]])[0][:, 0, 0]
    np.testing.assert_allclose(a * c[s] ** 2, b * c[r] ** 2, rtol=1e-9, atol=1e-12 * np.abs(a).max() * c[s] ** 2)


def _grad_case(seed=3):
    nz, nx, h, dt = 26, 30, 10.0, 1e-3
    c = _model(nz, nx, seed)
    d = coeffs.pml_coefficients_2d(nz, nx, N=6)
    w = coeffs.ricker(15.0, dt, 90, delay=20)
    src = [[8, 7], [8, 22]]
    rec = [[[7, x] for x in range(3, nx - 3)], [[7, x] for x in range(2, nx - 2, 2)]]
    c_true = c.copy(); c_true[12:18, 10:20] *= 1.05
    obs = oa.forward(c_true, d, dt, h, w, src, rec)
    return c, d, dt, h, w, src, rec, obs


def test_gradient_finite_difference():
    c, d, dt, h, w, src, rec, obs = _grad_case()
    J, g, _ = oa.gradient(c, d, dt, h, w, src, rec, obs)
    rng = np.random.default_rng(7)
    cells = [(int(rng.integers(2, c.shape[0] - 2)), int(rng.integers(2, c.shape[1] - 2))) for _ in range(5)]
    for (i, j) in cells:
        eps = 1e-5 * c[i, j]
        cp = c.copy(); cp[i, j] += eps
        cm = c.copy(); cm[i, j] -= eps
        fd = (oa.l2(oa.forward(cp, d, dt, h, w, src, rec), obs)
              - oa.l2(oa.forward(cm, d, dt, h, w, src, rec), obs)) / (2 * eps)
        assert abs(fd - g[i, j]) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(g).max()), (i, j, fd, g[i, j])


def test_dot_product():
    c, d, dt, h, w, src, rec, _ = _grad_case(4)
    rng = np.random.default_rng(11)
    dc = rng.standard_normal(c.shape)
    dd = [rng.standard_normal((w.shape[0], len(r), 1)) for r in rec]
    syn = oa.forward(c, d, dt, h, w, src, rec)
    obs = [s - 0.5 * x for s, x in zip(syn, dd)]  # residual 2(syn-obs) == dd
    _, g, _ = oa.gradient(c, d, dt, h, w, src, rec, obs)
    lhs = sum(np.sum(a * b) for a, b in zip(oa.born_apply(c, d, dt, h, w, src, rec, dc), dd))
    rhs = np.sum(dc * g)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_gradient_matches_torch_autograd_fp64():
    c, d, dt, h, w, src, rec, obs = _grad_case(5)
    J, g, syn = oa.gradient(c, d, dt, h, w, src, rec, obs)
    Jt, gt, synt = ref_torch.gradient_ad(c, d, dt, h, w, src, rec, obs, dtype=torch.float64)
    np.testing.assert_allclose(float(Jt), J, rtol=1e-10)
    np.testing.assert_allclose(gt.numpy(), g, rtol=1e-8, atol=1e-10 * np.abs(g).max())
    for a, b in zip(syn, synt):
        np.testing.assert_allclose(b.numpy(), a, rtol=1e-10, atol=1e-14)


@pytest.mark.parametrize("dtype", [torch.float64])
def test_boundary_saving_matches_ad_on_interior(dtype):
    nz, nx, h, dt = 112, 118, 10.0, 1e-3
    c = _model(nz, nx, 6)
    d = coeffs.pml_coefficients_2d(nz, nx, N=50)
    w = coeffs.ricker(15.0, dt, 120, delay=30)
    src = [[52, 58], [53, 61]]
    rec = [[[50, x] for x in range(50, nx - 50)]] * 2
    c_true = c.copy(); c_true[55:60, 55:62] *= 1.05
    obs = oa.forward(c_true, d, dt, h, w, src, rec)
    J, g, _ = oa.gradient(c, d, dt, h, w, src, rec, obs)
    Jb, gb, _ = ref_torch.gradient_bs(c, d, dt, h, w, src, rec, obs, dtype=dtype)
    np.testing.assert_allclose(float(Jb), J, rtol=1e-10)
    gi = g[50:-50, 50:-50]
    np.testing.assert_allclose(gb.numpy()[50:-50, 50:-50], gi, rtol=1e-7, atol=1e-9 * np.abs(gi).max())


@pytest.mark.parametrize("habc", [False, True])
def test_boundary_saving_free_surface_matches_ad(habc):
    """The boundary-saving restatement with a free surface (ring without the
    top row, rows rebuilt from row 0) equals FP64 autograd through the loop on
    that region, for the PML and the HABC step; zero elsewhere."""
    nz, nx, h, dt, P = 62, 118, 10.0, 1e-3, 50
    c = _model(nz, nx, 8)
    if habc:
        d = coeffs.habc_coefficients_2d(nz, nx, w=P, multiple=True)
    else:
        d = coeffs.pml_coefficients_2d(nz, nx, N=P, multiple=True)
    w = coeffs.ricker(15.0, dt, 100, delay=30)
    src = [[2, 58], [3, 61]]
    rec = [[[0, x] for x in range(50, nx - 50)]] * 2
    G = np.random.default_rng(1).standard_normal((100, 2 * (nx - 100)))
    if habc:
        ga, _ = ref_torch.habc_gradient_ad(c, d, dt, h, w, src, rec, G, dtype=torch.float64,
                                           multiple=True)
    else:
        _, ga, _ = ref_torch.gradient_ad(c, d, dt, h, w, src, rec, None, dtype=torch.float64,
                                         rec_cotangent=G)
    _, gb, _ = ref_torch.gradient_bs(c, d, dt, h, w, src, rec, None, dtype=torch.float64,
                                     multiple=True, rec_cotangent=G, habc=habc)
    region = (slice(0, nz - P), slice(P, nx - P))
    gi = ga.numpy()[region]
    np.testing.assert_allclose(gb.numpy()[region], gi, rtol=1e-7, atol=1e-9 * np.abs(gi).max())
    assert np.abs(gi[:5]).max() > 1e-3 * np.abs(gi).max()
    mask = np.ones((nz, nx), bool)
    mask[region] = False
    assert np.all(gb.numpy()[mask] == 0.0)
