"""FP64 oracle for the acoustic HABC step (per-cell formulation).

TEST INFRASTRUCTURE: imported only by tests/, smoke() and bench.py's
cpu_baseline leg; never by the product.

Semantics restated from seistorch/equations2d/acoustic_habc.py:79-221 and
seistorch/habc.py:4-81, written cell by cell (the reference works on rotated
51-row strips; oracle/ref_torch.habc_step restates it that way, and the two
are checked against each other):

* y = undamped update 2u - u_prev + c^2 dt^2 L u written as
  (1/dt^-2)(2/dt^2 u - dt^-2 u_prev + c^2 L u)            (:206-217)
* each boundary strip s (top, bottom, left, right; 50 cells) has its
  distance i from the outer edge and inward unit step e; its one-way value
  u1_s = (2-l-m) u + (l+2m) u[+e] - m u[+2e] + (l-1) p + (-l) p[+e]
  with l = 2c dt/h, m = c^2 dt^2/h^2 (:88-97), and
  y_s = b u1_s + (1-b) y with b the HABC weight of the cell (:99);
* the value is y, replaced in order by y_top, y_bottom, y_left, y_right
  wherever that strip's trapezoid mask holds (:167-183), then the four
  diagonals are averaged 0.5 y_s1 + 0.5 y_s2 (:185-202); with a free surface
  (multiple) the top strip and the two top diagonals are skipped and the
  side masks cover the top 50 rows fully (habc.py:224-235).
The reference extracts each strip as 51 rows and shifts them with
torch.roll, so in strip row 49 the "two cells inward" term wraps around to the
strip's row 0 (its outer edge).  Where the blend weight of such a cell is
nonzero (the far ends of the top-right and bottom-left diagonals, whose weight
comes from the other strip) that wrap is visible, so the oracle reproduces it.
"""
import numpy as np

W = 50


def masks(nz, nx, multiple=False, w=W):
    """Boolean (nz, nx) masks of the four strips as applied to the field."""
    z = np.arange(nz)[:, None]
    x = np.arange(nx)[None, :]
    top = (z < w) & (x >= z) & (x <= nx - 1 - z)
    ib = nz - 1 - z
    bottom = (z >= nz - w) & (x >= ib) & (x <= nx - 1 - ib)
    left = (x < w) & (z >= x) & (z <= nz - 1 - x)
    ir = nx - 1 - x
    right = (x >= nx - w) & (z >= ir) & (z <= nz - 1 - ir)
    if multiple:
        top = np.zeros_like(top)
        left = left | ((x < w) & (z < w))
        right = right | ((x >= nx - w) & (z < w))
    return top, bottom, left, right


def one_way(u, p, c, dt, h, axis, outer_low):
    """u1 for the strip on the low (outer_low) or high side of `axis`."""
    lam = 2 * c * dt / h
    mu = c ** 2 * dt ** 2 / h ** 2
    step = 1 if outer_low else -1

    def sh(f, k):
        # value k cells inward; cells whose inward neighbours leave the grid get 0
        out = np.zeros_like(f)
        n = f.shape[axis]
        src = [slice(None)] * f.ndim
        dst = [slice(None)] * f.ndim
        if step > 0:
            src[axis] = slice(k, n)
            dst[axis] = slice(0, n - k)
        else:
            src[axis] = slice(0, n - k)
            dst[axis] = slice(k, n)
        out[tuple(dst)] = f[tuple(src)]
        return out

    u2 = sh(u, 2)
    # strip row 49 (distance W-1 from the edge): the rolled 51-row strip wraps
    # its "+2" neighbour to row 0, the edge cell of the same line
    n = u.shape[axis]
    r49 = [slice(None)] * u.ndim
    r0 = [slice(None)] * u.ndim
    r49[axis] = W - 1 if outer_low else n - W
    r0[axis] = 0 if outer_low else n - 1
    u2[tuple(r49)] = u[tuple(r0)]
    return ((2 - lam - mu) * u + (lam + 2 * mu) * sh(u, 1) - mu * u2
            + (lam - 1) * p - lam * sh(p, 1))


def step(c, u, p, b, dt, h, multiple=False):
    """One HABC time step on (B, nz, nx) fields."""
    from .acoustic import laplacian
    c = np.asarray(c, np.float64)
    b = np.asarray(b, np.float64)
    u = np.asarray(u, np.float64)
    p = np.asarray(p, np.float64)
    nz, nx = c.shape
    y = (1.0 / dt ** -2) * (2 / dt ** 2 * u - dt ** -2 * p + c ** 2 * laplacian(u, h))
    ys = {"top": one_way(u, p, c, dt, h, -2, True),
          "bottom": one_way(u, p, c, dt, h, -2, False),
          "left": one_way(u, p, c, dt, h, -1, True),
          "right": one_way(u, p, c, dt, h, -1, False)}
    ys = {k: b * v + (1 - b) * y for k, v in ys.items()}
    top, bottom, left, right = masks(nz, nx, multiple)
    out = y.copy()
    for name, m in (("top", top), ("bottom", bottom), ("left", left), ("right", right)):
        out[:, m] = ys[name][:, m]
    i = np.arange(W)
    diag = []
    if not multiple:
        diag += [((i, i), "top", "left"), ((i, nx - W + i), "top", "right")]
    diag += [((nz - W + i, i), "bottom", "left"), ((nz - W + i, nx - W + i), "bottom", "right")]
    for (zz, xx), s1, s2 in diag:
        out[:, zz, xx] = 0.5 * ys[s1][:, zz, xx] + 0.5 * ys[s2][:, zz, xx]
    return out


def forward(c, b, dt, h, wavelet, src, rec, multiple=False, keep_fields=False):
    """rnn.py:178-212 loop with the HABC step; records (nt, nrec, 1) per shot."""
    c = np.asarray(c, np.float64)
    w = np.asarray(wavelet, np.float64)
    src = np.asarray(src, np.int64).reshape(-1, 2)
    recs = [np.asarray(r, np.int64).reshape(-1, 2) for r in rec]
    B = src.shape[0]
    nz, nx = c.shape
    u = np.zeros((B, nz, nx))
    p = np.zeros_like(u)
    out = [np.zeros((w.shape[0], r.shape[0], 1)) for r in recs]
    fields = [u.copy()] if keep_fields else None
    bidx = np.arange(B)
    for n in range(w.shape[0]):
        y = step(c, u, p, b, dt, h, multiple)
        y[bidx, src[:, 0], src[:, 1]] += w[n]
        p, u = u, y
        for k in range(B):
            out[k][n, :, 0] = u[k, recs[k][:, 0], recs[k][:, 1]]
        if keep_fields:
            fields.append(u.copy())
    return (out, fields) if keep_fields else out
