"""Pins the HABC oracle.

* two independent restatements agree: the FP64 per-cell oracle
  (oracle/habc.py) and the strip-based torch restatement (oracle/ref_torch,
  rolls/flips/transposes as in the reference) run in FP64;
* closed-form boundary behaviour: a plane wave leaving the grid through the
  HABC strips is absorbed far better than by the bare zero-padded step;
* interior cells (>= 50 from every edge) follow the undamped step exactly.
"""
import numpy as np
import pytest
import torch

from oracle import acoustic as oa
from oracle import coeffs
from oracle import habc as oh
from oracle import ref_torch


def _fields(nz, nx, seed):
    rng = np.random.default_rng(seed)
    c = 1500 + 1000 * rng.random((nz, nx))
    u = rng.standard_normal((2, nz, nx))
    p = rng.standard_normal((2, nz, nx))
    return c, u, p


@pytest.mark.parametrize("multiple", [False, True])
def test_per_cell_oracle_matches_strip_restatement(multiple):
    nz, nx, dt, h = 130, 150, 1e-3, 10.0
    c, u, p = _fields(nz, nx, 1)
    b = coeffs.habc_coefficients_2d(nz, nx, 50, multiple)
    o = oh.step(c, u, p, b, dt, h, multiple)
    t, _ = ref_torch.habc_step(torch.from_numpy(c), torch.from_numpy(u), torch.from_numpy(p),
                               torch.tensor(dt, dtype=torch.float64),
                               torch.tensor(h, dtype=torch.float64), torch.from_numpy(b),
                               multiple=multiple)
    np.testing.assert_allclose(t.numpy(), o, rtol=1e-12, atol=1e-12 * np.abs(o).max())


def test_interior_is_the_undamped_step():
    nz, nx, dt, h = 130, 150, 1e-3, 10.0
    c, u, p = _fields(nz, nx, 2)
    b = coeffs.habc_coefficients_2d(nz, nx)
    o = oh.step(c, u, p, b, dt, h)
    y = oa.step(c, u, p, np.zeros_like(c), dt, h)
    np.testing.assert_allclose(o[:, 50:-50, 50:-50], y[:, 50:-50, 50:-50], rtol=1e-12)
    # the 49th ring has weight 0: also the undamped value there
    np.testing.assert_allclose(o[:, 49, 60:-60], y[:, 49, 60:-60], rtol=1e-12)


def test_habc_absorbs_outgoing_waves():
    """A point source in a 160x160 homogeneous grid: the bare zero-padded step
    reflects the outgoing wave back into the interior (energy returns after
    ~450 steps), the HABC strips absorb it."""
    nz, nx, dt, h = 160, 160, 1e-3, 10.0
    c = np.full((nz, nx), 3000.0)
    w = coeffs.ricker(15.0, dt, 700, delay=60)
    src = [[80, 80]]
    rec = [[[80, 80]]]
    b = coeffs.habc_coefficients_2d(nz, nx)
    _, f_habc = oh.forward(c, b, dt, h, w, src, rec, keep_fields=True)
    _, f_bare = oa.forward(c, np.zeros((nz, nx)), dt, h, w, src, rec, keep_fields=True)
    energy = lambda f: float(np.sum(f[0, 50:-50, 50:-50] ** 2))  # noqa: E731
    peak = max(energy(f) for f in f_habc[:200])
    e_h, e_b = energy(f_habc[-1]), energy(f_bare[-1])
    assert e_b > 0.05 * peak          # the bare boundary reflects
    assert e_h < 1e-5 * peak and e_h < 1e-4 * e_b, (e_h, e_b, peak)
