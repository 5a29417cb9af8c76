"""FP64 restatements of the boundary-coefficient and wavelet generators.

TEST INFRASTRUCTURE: imported only by tests/, smoke() and bench.py's
cpu_baseline leg.

Pinned by closed-form values in tests/test_oracle_coeffs.py:
* PML: d0 = 1.5*1500/50*log10(1e4) = 180, d[0] = 180, d[49] = 0.072,
  d[50] = 0, corner = sqrt(2)*180, separability d^2 = px^2 + pz^2.
* HABC: weight row i = (49-i)/49, trapezoid row r keeps columns r..nx-1-r.
* Ricker: peak 1 at i == delay, zero crossings at c = +-1/sqrt(2),
  closed form of its spectrum is not needed.
"""
import numpy as np


def pml_profile(N=50, cp=1500.0, R=1e-4, order=2):
    """1-D PML damping profile, outermost cell first.

    seistorch/pml.py:35-38: d0 = (1.5*cp/N)*log10(1/R);
    d_vals = d0*linspace(0,1,N+1)**order, flipped so index 0 is the edge.
    """
    d0 = (1.5 * cp / N) * np.log10(1.0 / R)
    vals = d0 * np.linspace(0.0, 1.0, N + 1) ** order
    return vals[::-1].copy()


def pml_axis(n, N=50, skip_low=False):
    """Damping along one axis of length n (both ends, or only the high end)."""
    prof = pml_profile(N)
    p = np.zeros(n)
    if N > 0:
        if not skip_low:
            p[: N + 1] = prof
        p[n - N - 1:] = prof[::-1]
    return p


def pml_coefficients_2d(nz, nx, N=50, multiple=False):
    """d(z, x) = sqrt(pz(z)^2 + px(x)^2); top strip omitted when multiple.

    seistorch/pml.py:29-59 (domain_shape=(nz, nx) is unpacked as Nx, Ny and
    transposed back at :56, so x and z both get the same profile).
    """
    pz = pml_axis(nz, N, skip_low=multiple)
    px = pml_axis(nx, N)
    return np.sqrt(pz[:, None] ** 2 + px[None, :] ** 2)


def habc_masks(nz, nx, w=50):
    """Trapezoid strip masks (top, bottom, left, right) of seistorch/habc.py:4-40.

    top[i, j] = 1 iff i <= j <= nx-1-i  (i < w); bottom = flipud(top);
    left[i, j] = 1 iff j <= i and j <= nz-1-i (j < w); right = fliplr(left).
    """
    i = np.arange(w)[:, None]
    j = np.arange(nx)[None, :]
    top = ((j >= i) & (j <= nx - 1 - i)).astype(np.float64)
    bottom = top[::-1].copy()
    ii = np.arange(nz)[:, None]
    jj = np.arange(w)[None, :]
    left = ((jj <= ii) & (jj <= nz - 1 - ii)).astype(np.float64)
    right = left[:, ::-1].copy()
    return top, bottom, left, right


def habc_weights(w=50):
    """Blend weight per strip row, row 0 at the outer edge.

    seistorch/habc.py:51-52: linspace(0, N, N)/N flipped, i.e. (N-1-i)/(N-1).
    """
    return np.linspace(0.0, w, w)[::-1] / w


def habc_coefficients_2d(nz, nx, w=50, multiple=False):
    """Weight field of seistorch/habc.py:42-81 (trapezoid-masked strips)."""
    d = np.zeros((nz, nx))
    b = habc_weights(w)
    top, bottom, left, right = habc_masks(nz, nx, w)
    if not multiple:
        d[:w][top == 1] = np.repeat(b[:, None], nx, 1)[top == 1]
    d[-w:][bottom == 1] = np.repeat(b[::-1][:, None], nx, 1)[bottom == 1]
    lm = left.copy()
    rm = right.copy()
    if multiple:
        lm[:w] = 1.0
        rm[:w] = 1.0
    sub = d[:, :w]
    sub[lm == 1] = np.repeat(b[None, :], nz, 0)[lm == 1]
    sub = d[:, -w:]
    sub[rm == 1] = np.repeat(b[::-1][None, :], nz, 0)[rm == 1]
    return d


def ricker(fm, dt, nt, delay=80, inverse=False):
    """Ricker wavelet r_i = (1-2c^2) exp(-c^2), c = pi*fm*(i*dt - delay*dt).

    seistorch/utils.py:235-249 (delay is given in samples).
    """
    i = np.arange(nt)
    c = np.pi * fm * (i * dt - delay * dt)
    p = -1.0 if inverse else 1.0
    return p * (1 - 2 * c ** 2) * np.exp(-c ** 2)
