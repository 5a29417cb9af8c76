"""FP64 oracle for the constant-density acoustic PML path.

TEST INFRASTRUCTURE: imported only by tests/, smoke() and bench.py's
cpu_baseline leg; never by the product.

Semantics restated from the reference:
* step      seistorch/equations2d/acoustic.py:73-86
              y = A*(2/dt^2*u - B*u_prev + c^2*L u),
              A = 1/(1/dt^2 + d/dt), B = 1/dt^2 - d/dt,
              L = zero-padded 5-point Laplacian / h^2 (acoustic.py:5-71; the
              Taylor solve for spatial order 2 gives taps [1, -4, 1]).
* loop      seistorch/rnn.py:178-212: fields start at zero; per step
              (1) step, (2) raw add of w[n] at the shot's source cell
              (source.py:47-52, smask built at rnn.py:158-166, no dt^2/v^2
              scaling), (3) record receivers from the post-source field
              (probe.py:42-44). Records are (nt, nrec, 1) per shot.
* misfit    seistorch/loss.py:409-421: J = sum over shots of sum (syn-obs)^2.

The adjoint is the hand-derived exact discrete transpose of that loop
(mu = A*lambda):
    mu^k  = A*(s^k + 2/dt^2*mu^{k+1} - B*mu^{k+2} + L(c^2*mu^{k+1}))
    dJ/dc = sum_n 2c * (L u^n) * mu^{n+1}
with s^k = 2(r[k-1]-obs[k-1]) at the receivers.  It is pinned by central
finite differences of J and by torch.autograd of the same loop
(tests/test_oracle_acoustic.py); the forward is pinned by the exact
standing-wave solution of the leapfrog scheme and by the closed-form
two-root recursion of a single damped cell.
"""
import numpy as np


def laplacian(u, h):
    """Zero-padded 5-point Laplacian over the last two axes (acoustic.py:60-71)."""
    out = -4.0 * u
    out[..., 1:, :] += u[..., :-1, :]
    out[..., :-1, :] += u[..., 1:, :]
    out[..., :, 1:] += u[..., :, :-1]
    out[..., :, :-1] += u[..., :, 1:]
    return out / (h * h)


def damping_ab(d, dt):
    """A and B of acoustic.py:81-82."""
    A = 1.0 / (1.0 / dt ** 2 + d / dt)
    B = 1.0 / dt ** 2 - d / dt
    return A, B


def step(c, u, up, d, dt, h):
    """One PML time step without source (acoustic.py:73-86)."""
    A, B = damping_ab(d, dt)
    return A * (2.0 / dt ** 2 * u - B * up + c ** 2 * laplacian(u, h))


def _as_shots(src, rec):
    src = np.asarray(src, dtype=np.int64).reshape(-1, 2)
    recs = [np.asarray(r, dtype=np.int64).reshape(-1, 2) for r in rec]
    assert len(recs) == src.shape[0]
    return src, recs


def forward(c, d, dt, h, wavelet, src, rec, keep_fields=False):
    """Batched forward modelling.

    c, d      : (nz, nx) velocity and PML damping on the padded grid.
    wavelet   : (nt,) samples, shared by every shot (rnn.py:187,198).
    src       : (B, 2) padded (z, x) source cell per shot.
    rec       : list of B arrays (nrec_b, 2) padded (z, x) receiver cells.
    returns   : list of B records (nt, nrec_b, 1) [, fields u^0..u^nt].
    """
    c = np.asarray(c, np.float64)
    d = np.asarray(d, np.float64)
    w = np.asarray(wavelet, np.float64)
    src, recs = _as_shots(src, rec)
    B = src.shape[0]
    nz, nx = c.shape
    nt = w.shape[0]
    u = np.zeros((B, nz, nx))
    up = np.zeros((B, nz, nx))
    out = [np.zeros((nt, r.shape[0], 1)) for r in recs]
    fields = [u.copy()] if keep_fields else None
    bidx = np.arange(B)
    for n in range(nt):
        y = step(c, u, up, d, dt, h)
        y[bidx, src[:, 0], src[:, 1]] += w[n]
        up, u = u, y
        for b in range(B):
            out[b][n, :, 0] = u[b, recs[b][:, 0], recs[b][:, 1]]
        if keep_fields:
            fields.append(u.copy())
    if keep_fields:
        return out, fields
    return out


def l2(syn, obs):
    """sum_b sum (syn_b - obs_b)^2 (loss.py:409-421)."""
    return float(sum(np.sum((np.asarray(s, np.float64) - np.asarray(o, np.float64)) ** 2)
                     for s, o in zip(syn, obs)))


def gradient(c, d, dt, h, wavelet, src, rec, obs):
    """Exact discrete adjoint: returns (J, dJ/dc, records).

    Stores every forward field, so it is meant for small grids only.
    """
    c = np.asarray(c, np.float64)
    d = np.asarray(d, np.float64)
    src, recs = _as_shots(src, rec)
    syn, fields = forward(c, d, dt, h, wavelet, src, recs, keep_fields=True)
    J = l2(syn, obs)
    B = src.shape[0]
    nt = len(fields) - 1
    A, Bc = damping_ab(d, dt)
    c2 = c * c
    mu1 = np.zeros_like(fields[0])  # mu^{k+1}
    mu2 = np.zeros_like(fields[0])  # mu^{k+2}
    g = np.zeros_like(c)
    for k in range(nt, 0, -1):
        lam = 2.0 / dt ** 2 * mu1 - Bc * mu2 + laplacian(c2 * mu1, h)
        for b in range(B):
            res = 2.0 * (syn[b][k - 1, :, 0] - np.asarray(obs[b], np.float64)[k - 1, :, 0])
            np.add.at(lam[b], (recs[b][:, 0], recs[b][:, 1]), res)
        mu = A * lam
        g += np.sum(2.0 * c * laplacian(fields[k - 1], h) * mu, axis=0)
        mu2, mu1 = mu1, mu
    return J, g, syn


def born_apply(c, d, dt, h, wavelet, src, rec, dc):
    """Linearised records dr = (dr/dc) dc (forward-mode), for dot-product tests."""
    c = np.asarray(c, np.float64)
    d = np.asarray(d, np.float64)
    dc = np.asarray(dc, np.float64)
    src, recs = _as_shots(src, rec)
    w = np.asarray(wavelet, np.float64)
    B = src.shape[0]
    nz, nx = c.shape
    A, _ = damping_ab(d, dt)
    u = np.zeros((B, nz, nx)); up = np.zeros_like(u)
    du = np.zeros_like(u); dup = np.zeros_like(u)
    out = [np.zeros((w.shape[0], r.shape[0], 1)) for r in recs]
    bidx = np.arange(B)
    for n in range(w.shape[0]):
        Lu = laplacian(u, h)
        y = step(c, u, up, d, dt, h)
        dy = step(c, du, dup, d, dt, h) + A * 2.0 * c * dc * Lu
        y[bidx, src[:, 0], src[:, 1]] += w[n]
        up, u = u, y
        dup, du = du, dy
        for b in range(B):
            out[b][n, :, 0] = du[b, recs[b][:, 0], recs[b][:, 1]]
    return out


def ring_cells(nz, nx, npml=49):
    """(z, x) cells of the 1-cell ring that boundary saving stores.

    seistorch/equations2d/utils.py:50-69 saves rows npml and -(npml+1) and
    columns npml and -(npml+1); the reconstruction of the undamped interior
    [npml+1, n-npml-1) only reads the ring cells that are 5-point neighbours
    of that interior, i.e. the ring without its four corners.
    """
    z0, z1 = npml, nz - 1 - npml
    x0, x1 = npml, nx - 1 - npml
    xs = np.arange(x0 + 1, x1)
    zs = np.arange(z0 + 1, z1)
    cells = [np.stack([np.full_like(xs, z0), xs], 1),
             np.stack([np.full_like(xs, z1), xs], 1),
             np.stack([zs, np.full_like(zs, x0)], 1),
             np.stack([zs, np.full_like(zs, x1)], 1)]
    return np.concatenate(cells, 0)
