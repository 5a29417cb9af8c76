"""Pins the FP64 elastic oracle without restating its formulas.

* P- and S-wave travel times in a homogeneous medium (explosive source for P,
  shear-stress source for S) against the closed-form speeds;
* causality of the radius-1 staggered stencil, zero source, linearity;
* the adjoint: central finite differences on vp, vs and rho, the dot-product
  test, and torch autograd (FP64) of the op-for-op torch restatement.
"""
import numpy as np
import torch

from oracle import coeffs
from oracle import elastic as oe
from oracle import ref_torch


def _homog(nz, nx, vp=2000.0, vs=1150.0, rho=2000.0):
    return (np.full((nz, nx), vp), np.full((nz, nx), vs), np.full((nz, nx), rho))


def _peak_time(trace, dt):
    return np.argmax(np.abs(trace)) * dt


def test_p_and_s_travel_times():
    nz, nx, h, dt, nt = 140, 140, 5.0, 5e-4, 700
    vp, vs, rho = _homog(nz, nx)
    d = coeffs.pml_coefficients_2d(nz, nx, N=20)
    fm, delay = 15.0, 100
    w = coeffs.ricker(fm, dt, nt, delay=delay)
    src = [[70, 70]]
    dist = 50
    rec = [[[70, 70 + dist]]]
    # explosive source: a P wave; vx at a receiver along x peaks at r/vp after the source peak
    r = oe.forward(vp, vs, rho, d, dt, h, w, src, rec, ("txx", "tzz"), ("vx",))[0][:, 0, 0]
    tp = _peak_time(r, dt) - delay * dt
    assert abs(tp - dist * h / 2000.0) < 0.05 * dist * h / 2000.0, tp
    # shear stress source: the vz wave along x travels at vs
    r = oe.forward(vp, vs, rho, d, dt, h, w, src, rec, ("txz",), ("vz",))[0][:, 0, 0]
    ts = _peak_time(r, dt) - delay * dt
    assert abs(ts - dist * h / 1150.0) < 0.05 * dist * h / 1150.0, ts


def test_causality_zero_source_linearity():
    nz, nx, h, dt = 30, 34, 5.0, 5e-4
    vp, vs, rho = _homog(nz, nx)
    d = np.zeros((nz, nx))
    w = np.zeros(8); w[0] = 1.0
    src = [[15, 17]]
    rec = [[[15, 17 + k] for k in range(8)]]
    r = oe.forward(vp, vs, rho, d, dt, h, w, src, rec, ("txx", "tzz"), ("vx",))[0][:, :, 0]
    # the impulse enters after step 0 and moves <= 1 cell per half step pair
    for k in range(2, 8):
        assert np.all(r[: k - 1, k] == 0.0), k
    assert np.any(r[:, 2] != 0)
    w2 = coeffs.ricker(20.0, dt, 60, delay=20)
    a = oe.forward(vp, vs, rho, d, dt, h, w2, src, rec)[0]
    b = oe.forward(vp, vs, rho, d, dt, h, -3.0 * w2, src, rec)[0]
    np.testing.assert_allclose(b, -3.0 * a, rtol=1e-12, atol=1e-12 * np.abs(a).max())
    z = oe.forward(vp, vs, rho, d, dt, h, 0 * w2, src, rec)[0]
    assert np.all(z == 0)


def _case(seed=0):
    rng = np.random.default_rng(seed)
    nz, nx, h, dt, nt = 24, 28, 5.0, 4e-4, 80
    vp = 2000 + 200 * rng.random((nz, nx))
    vs = vp / 1.73 * (1 + 0.05 * rng.random((nz, nx)))
    vs[:4] = 0.0  # water layer
    rho = 310 * vp ** 0.25
    d = coeffs.pml_coefficients_2d(nz, nx, N=5)
    w = coeffs.ricker(25.0, dt, nt, delay=20)
    src = [[7, 8], [7, 19]]
    rec = [[[5, x] for x in range(3, nx - 3)]] * 2
    return vp, vs, rho, d, dt, h, w, src, rec


def test_gradient_finite_difference():
    vp, vs, rho, d, dt, h, w, src, rec = _case(1)
    obs = oe.forward(vp * 1.02, vs, rho * 0.99, d, dt, h, w, src, rec)
    J, (gvp, gvs, grho), _ = oe.gradient(vp, vs, rho, d, dt, h, w, src, rec, obs)
    rng = np.random.default_rng(2)
    for which, g in (("vp", gvp), ("vs", gvs), ("rho", grho)):
        for _ in range(3):
            i = int(rng.integers(6, 20)); j = int(rng.integers(4, 24))
            vals = []
            for sgn in (1, -1):
                m = {"vp": vp.copy(), "vs": vs.copy(), "rho": rho.copy()}
                eps = 1e-6 * m[which][i, j]
                m[which][i, j] += sgn * eps
                syn = oe.forward(m["vp"], m["vs"], m["rho"], d, dt, h, w, src, rec)
                vals.append(sum(np.sum((s - o) ** 2) for s, o in zip(syn, obs)))
            fd = (vals[0] - vals[1]) / (2 * 1e-6 * (vp if which == "vp" else vs if which == "vs" else rho)[i, j])
            assert abs(fd - g[i, j]) <= 1e-5 * max(abs(fd), 1e-4 * np.abs(g).max()), (which, i, j, fd, g[i, j])


def test_gradient_matches_torch_autograd_fp64():
    vp, vs, rho, d, dt, h, w, src, rec = _case(3)
    obs = oe.forward(vp * 1.01, vs * 0.98, rho, d, dt, h, w, src, rec)
    J, g, syn = oe.gradient(vp, vs, rho, d, dt, h, w, src, rec, obs)
    Jt, gt, synt = ref_torch.elastic_gradient_ad(vp, vs, rho, d, dt, h, w, src, rec,
                                                 [torch.from_numpy(o) for o in obs],
                                                 dtype=torch.float64)
    np.testing.assert_allclose(float(Jt), J, rtol=1e-10)
    for a, b in zip(g, gt):
        np.testing.assert_allclose(b.numpy(), a, rtol=1e-8, atol=1e-10 * np.abs(a).max())
    for a, b in zip(syn, synt):
        np.testing.assert_allclose(b.numpy(), a, rtol=1e-10, atol=1e-14)


def test_dot_product():
    vp, vs, rho, d, dt, h, w, src, rec = _case(4)
    rng = np.random.default_rng(5)
    dm = [rng.standard_normal(vp.shape) * s for s in (10.0, 10.0, 5.0)]
    eps = 1e-6
    syn0 = oe.forward(vp, vs, rho, d, dt, h, w, src, rec)
    dd = [rng.standard_normal(s.shape) for s in syn0]
    obs = [s - 0.5 * x for s, x in zip(syn0, dd)]
    _, g, _ = oe.gradient(vp, vs, rho, d, dt, h, w, src, rec, obs)
    sp = oe.forward(vp + eps * dm[0], vs + eps * dm[1], rho + eps * dm[2], d, dt, h, w, src, rec)
    sm = oe.forward(vp - eps * dm[0], vs - eps * dm[1], rho - eps * dm[2], d, dt, h, w, src, rec)
    lhs = sum(np.sum((a - b) / (2 * eps) * x) for a, b, x in zip(sp, sm, dd))
    rhs = sum(np.sum(a * b) for a, b in zip(dm, g))
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
