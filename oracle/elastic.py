"""FP64 oracle for the 2D isotropic elastic velocity-stress path.

TEST INFRASTRUCTURE: imported only by tests/, smoke() and bench.py's
cpu_baseline leg; never by the product.

Semantics restated from the reference:
* differences  seistorch/equations2d/utils.py:3-48 (diff_using_roll):
               "forward" D-f[i] = f[i] - f[i-1], 0 at i = 0;
               "backward" D+f[i] = f[i+1] - f[i], 0 at the last index;
               axis 1 of (B, nz, nx) is z, axis 2 is x.
* step         seistorch/equations2d/elastic.py:7-37: stresses first,
               lambda = rho (vp^2 - 2 vs^2), mu = rho vs^2, c = dt d / 2,
                 txx' = (txx(1-c) + dt/h((lam+2mu) Dx- vx + lam Dz+ vz)) / (1+c)
                 tzz' = (tzz(1-c) + dt/h((lam+2mu) Dz+ vz + lam Dx- vx)) / (1+c)
                 txz' = (txz(1-c) + dt mu/h (Dx+ vz + Dz- vx)) / (1+c)
               then velocities from the new stresses,
                 vx' = (vx(1-c) + dt/(rho h)(Dx+ txx' + Dz+ txz')) / (1+c)
                 vz' = (vz(1-c) + dt/(rho h)(Dx- txz' + Dz- tzz')) / (1+c)
* loop         seistorch/rnn.py:176-212 with time offset 0 for first-order
               equations: step, raw add of w[n] to each source field
               (source.py:47-52, no clone for first-order equations), record
               the receiver fields (probe.py:42-44), records (nt, nrec, nch).

The adjoint is the hand-derived exact transpose of that loop, with the model
gradients chained through P = dt/h rho vp^2, L = dt/h rho (vp^2 - 2 vs^2),
M = dt/h rho vs^2, R = dt/(h rho).  Pinned in tests/test_oracle_elastic.py by
central finite differences, the dot-product test and torch autograd (FP64)
of the op-for-op torch restatement; the forward by P- and S-wave travel times
in a homogeneous medium, causality of the radius-1 stencil, zero source and
linearity.
"""
import numpy as np

FIELDS = ("vx", "vz", "txx", "tzz", "txz")


def dminus(f, axis):
    """D-f[i] = f[i] - f[i-1], zero at index 0 (utils.py:5-23)."""
    out = np.zeros_like(f)
    sl = [slice(None)] * f.ndim
    sl_m = list(sl)
    sl[axis] = slice(1, None)
    sl_m[axis] = slice(None, -1)
    out[tuple(sl)] = f[tuple(sl)] - f[tuple(sl_m)]
    return out


def dplus(f, axis):
    """D+f[i] = f[i+1] - f[i], zero at the last index (utils.py:25-43)."""
    out = np.zeros_like(f)
    sl = [slice(None)] * f.ndim
    sl_p = list(sl)
    sl[axis] = slice(None, -1)
    sl_p[axis] = slice(1, None)
    out[tuple(sl)] = f[tuple(sl_p)] - f[tuple(sl)]
    return out


def dminus_T(g, axis):
    """Transpose of dminus: out[j] = [j>=1] g[j] - [j+1<=n-1] g[j+1]."""
    out = g.copy()
    sl0 = [slice(None)] * g.ndim
    sl0[axis] = 0
    out[tuple(sl0)] = 0.0
    a = [slice(None)] * g.ndim
    b = [slice(None)] * g.ndim
    a[axis] = slice(None, -1)
    b[axis] = slice(1, None)
    out[tuple(a)] -= g[tuple(b)]
    return out


def dplus_T(g, axis):
    """Transpose of dplus: out[j] = [j>=1] g[j-1] - [j<=n-2] g[j]."""
    out = -g.copy()
    sl = [slice(None)] * g.ndim
    sl[axis] = -1
    out[tuple(sl)] = 0.0
    a = [slice(None)] * g.ndim
    b = [slice(None)] * g.ndim
    a[axis] = slice(1, None)
    b[axis] = slice(None, -1)
    out[tuple(a)] += g[tuple(b)]
    return out


def coefficients(vp, vs, rho, d, dt, h):
    lam = rho * (vp ** 2 - 2 * vs ** 2)
    mu = rho * vs ** 2
    c = 0.5 * dt * d
    return dict(P=dt / h * (lam + 2 * mu), L=dt / h * lam, M=dt * mu / h, R=dt / (rho * h),
                a=1.0 / (1.0 + c), b=1.0 - c)


def step(k, vx, vz, txx, tzz, txz):
    """One velocity-stress step (elastic.py:7-37), returns the new fields and the
    intermediates the adjoint needs."""
    vx_x = dminus(vx, -1)
    vz_z = dplus(vz, -2)
    vx_z = dminus(vx, -2)
    vz_x = dplus(vz, -1)
    a, b = k["a"], k["b"]
    txx1 = a * (k["P"] * vx_x + k["L"] * vz_z + b * txx)
    tzz1 = a * (k["P"] * vz_z + k["L"] * vx_x + b * tzz)
    txz1 = a * (k["M"] * (vz_x + vx_z) + b * txz)
    sx = dplus(txx1, -1) + dplus(txz1, -2)
    sz = dminus(txz1, -1) + dminus(tzz1, -2)
    vx1 = a * (k["R"] * sx + b * vx)
    vz1 = a * (k["R"] * sz + b * vz)
    return (vx1, vz1, txx1, tzz1, txz1), (vx_x, vz_z, vz_x + vx_z, sx, sz)


def _setup(src, rec):
    src = np.asarray(src, np.int64).reshape(-1, 2)
    recs = [np.asarray(r, np.int64).reshape(-1, 2) for r in rec]
    return src, recs


def forward(vp, vs, rho, d, dt, h, wavelet, src, rec, src_type=("txx", "tzz"),
            rec_type=("vx", "vz"), keep=False):
    """Records: list of B arrays (nt, nrec_b, len(rec_type))."""
    k = coefficients(*(np.asarray(x, np.float64) for x in (vp, vs, rho, d)), dt, h)
    w = np.asarray(wavelet, np.float64)
    src, recs = _setup(src, rec)
    B = src.shape[0]
    nz, nx = np.shape(vp)
    f = {n: np.zeros((B, nz, nx)) for n in FIELDS}
    out = [np.zeros((w.shape[0], r.shape[0], len(rec_type))) for r in recs]
    hist = [] if keep else None
    bidx = np.arange(B)
    for n in range(w.shape[0]):
        new, inter = step(k, *(f[x] for x in FIELDS))
        if keep:
            hist.append(({x: f[x] for x in FIELDS}, inter))
        f = dict(zip(FIELDS, new))
        for st in src_type:
            f[st] = f[st].copy()
            f[st][bidx, src[:, 0], src[:, 1]] += w[n]
        for b in range(B):
            for ci, rt in enumerate(rec_type):
                out[b][n, :, ci] = f[rt][b, recs[b][:, 0], recs[b][:, 1]]
    return (out, hist, k) if keep else out


def gradient(vp, vs, rho, d, dt, h, wavelet, src, rec, obs, src_type=("txx", "tzz"),
             rec_type=("vx", "vz")):
    """Exact adjoint: (J, (g_vp, g_vs, g_rho), records); stores every step."""
    vp, vs, rho = (np.asarray(x, np.float64) for x in (vp, vs, rho))
    syn, hist, k = forward(vp, vs, rho, d, dt, h, wavelet, src, rec, src_type, rec_type, keep=True)
    src_a, recs = _setup(src, rec)
    J = float(sum(np.sum((s - np.asarray(o, np.float64)) ** 2) for s, o in zip(syn, obs)))
    B = src_a.shape[0]
    nz, nx = vp.shape
    adj = {n: np.zeros((B, nz, nx)) for n in FIELDS}
    gP = np.zeros((nz, nx)); gL = np.zeros((nz, nx)); gM = np.zeros((nz, nx)); gR = np.zeros((nz, nx))
    a, b = k["a"], k["b"]
    for n in range(len(hist) - 1, -1, -1):
        old, (vx_x, vz_z, shear, sx, sz) = hist[n]
        for bb in range(B):
            res = 2.0 * (syn[bb][n] - np.asarray(obs[bb], np.float64)[n])
            for ci, rt in enumerate(rec_type):
                np.add.at(adj[rt][bb], (recs[bb][:, 0], recs[bb][:, 1]), res[:, ci])
        # velocity update
        tx = a * k["R"] * adj["vx"]
        tz = a * k["R"] * adj["vz"]
        gR += np.sum(a * sx * adj["vx"] + a * sz * adj["vz"], 0)
        s_xx = adj["txx"] + dplus_T(tx, -1)
        s_xz = adj["txz"] + dplus_T(tx, -2) + dminus_T(tz, -1)
        s_zz = adj["tzz"] + dminus_T(tz, -2)
        vx_hat = a * b * adj["vx"]
        vz_hat = a * b * adj["vz"]
        # stress update
        u1, u2, u3 = a * s_xx, a * s_zz, a * s_xz
        vx_hat += dminus_T(k["P"] * u1 + k["L"] * u2, -1) + dminus_T(k["M"] * u3, -2)
        vz_hat += dplus_T(k["L"] * u1 + k["P"] * u2, -2) + dplus_T(k["M"] * u3, -1)
        gP += np.sum(vx_x * u1 + vz_z * u2, 0)
        gL += np.sum(vz_z * u1 + vx_x * u2, 0)
        gM += np.sum(shear * u3, 0)
        adj = {"vx": vx_hat, "vz": vz_hat, "txx": b * u1, "tzz": b * u2, "txz": b * u3}
    s = dt / h
    g_vp = gP * 2 * rho * vp * s + gL * 2 * rho * vp * s
    g_vs = gL * (-4) * rho * vs * s + gM * 2 * rho * vs * s
    g_rho = gP * vp ** 2 * s + gL * (vp ** 2 - 2 * vs ** 2) * s + gM * vs ** 2 * s \
        - gR * dt / (rho ** 2 * h)
    return J, (g_vp, g_vs, g_rho), syn
