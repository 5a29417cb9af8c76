"""Seeded synthetic models and acquisitions for the benchmark workloads.

There is no dataset access, so every workload is generated here from a seed
(20230503 by default, the reference's YAML default seed, template.yml:2).

* ``layered_model``      horizontal layers + Gaussian-smoothed random
                         perturbation (the 200x400 forward-modelling case).
* ``anomaly_model``      smooth vertical gradient + elliptical anomalies and
                         a Gaussian-smoothed starting model (the 800x2400
                         FWI-gradient case).
* ``surface_acquisition`` shots evenly spaced at a fixed depth and a receiver
                         at every surface grid point, in padded grid indices
                         (seistorch/setup.py:386-483 adds the boundary width).
"""
import numpy as np
from scipy.ndimage import gaussian_filter

DEFAULT_SEED = 20230503


def layered_model(nz, nx, vmin=1500.0, vmax=4000.0, nlayers=4, sigma=8.0, pert=0.05,
                  seed=DEFAULT_SEED):
    rng = np.random.default_rng(seed)
    v = np.empty((nz, nx), np.float64)
    speeds = np.linspace(vmin, vmax, nlayers)
    edges = np.linspace(0, nz, nlayers + 1).astype(int)
    for k in range(nlayers):
        v[edges[k]:edges[k + 1]] = speeds[k]
    noise = gaussian_filter(rng.standard_normal((nz, nx)), sigma)
    noise *= pert / max(np.abs(noise).max(), 1e-30)
    return (v * (1.0 + noise)).astype(np.float32)


def anomaly_model(nz, nx, vmin=1500.0, vmax=4500.0, n_anom=3, amp=0.10, smooth=20.0,
                  seed=DEFAULT_SEED):
    """(true, initial) models: smooth layered gradient with elliptical anomalies."""
    rng = np.random.default_rng(seed)
    z = np.linspace(0.0, 1.0, nz)[:, None]
    base = vmin + (vmax - vmin) * z ** 1.2 + 0.0 * np.arange(nx)[None, :]
    steps = gaussian_filter(np.repeat(np.floor(z * 6) / 6 * 0.08 * (vmax - vmin), nx, 1), 3.0)
    v = base + steps
    zz, xx = np.mgrid[0:nz, 0:nx]
    for _ in range(n_anom):
        cz = rng.uniform(0.25, 0.8) * nz
        cx = rng.uniform(0.15, 0.85) * nx
        az = rng.uniform(0.04, 0.1) * nz
        ax = rng.uniform(0.03, 0.08) * nx
        sign = rng.choice([-1.0, 1.0])
        inside = ((zz - cz) / az) ** 2 + ((xx - cx) / ax) ** 2 <= 1.0
        v = np.where(inside, v * (1.0 + sign * amp), v)
    true = v.astype(np.float32)
    init = gaussian_filter(v, smooth, mode="nearest").astype(np.float32)
    return true, init


def surface_acquisition(nx, nshots, src_depth=2, rec_depth=0, bwidth=50, first=None):
    """Source (z, x) per shot and the shared receiver line, padded indices."""
    if first is None:
        first = nx // (2 * nshots)
    xs = np.linspace(first, nx - 1 - first, nshots).round().astype(np.int64)
    src = np.stack([np.full(nshots, src_depth + bwidth), xs + bwidth], 1)
    rec = np.stack([np.full(nx, rec_depth + bwidth), np.arange(nx) + bwidth], 1)
    return src.astype(np.int64), [rec.astype(np.int64)] * nshots


def rec_flat_index(recs, nz, nx):
    """Flat (B, nz, nx) indices of each shot's receivers, shot-major.  recs[b]
    is an (n, 2) array of grid (z, x); a receiver outside the grid raises
    IndexError, as the reference's probe indexing does (probe.py:42-44)."""
    out = []
    for b, r in enumerate(recs):
        r = np.asarray(r, np.int64).reshape(-1, 2)
        bad = (r[:, 0] < 0) | (r[:, 0] >= nz) | (r[:, 1] < 0) | (r[:, 1] >= nx)
        if bad.any():
            z, x = r[np.argmax(bad)]
            raise IndexError(f"shot {b}: receiver (z={z}, x={x}) lies outside the "
                             f"{nz}x{nx} grid")
        out.append(b * nz * nx + r[:, 0] * nx + r[:, 1])
    return np.concatenate(out) if out else np.zeros(0, np.int64)


def elastic_models(nz, nx, vmin=1800.0, vmax=4500.0, water=20, pert=0.03, sigma=10.0,
                   seed=DEFAULT_SEED):
    """(vp, vs, rho): smooth layered vp with a seeded smooth perturbation,
    vs = vp/1.73 (0 in the top `water` rows), Gardner rho = 310 vp^0.25, each
    with its own seeded smooth perturbation."""
    rng = np.random.default_rng(seed)
    z = np.linspace(0.0, 1.0, nz)[:, None]
    vp = vmin + (vmax - vmin) * z + 0.0 * np.arange(nx)[None, :]
    vp = gaussian_filter(vp + np.floor(z * 5) / 5 * 0.1 * (vmax - vmin), 2.0)

    def smooth_noise():
        n = gaussian_filter(rng.standard_normal((nz, nx)), sigma)
        return n / max(np.abs(n).max(), 1e-30)

    vp = vp * (1.0 + pert * smooth_noise())
    vs = vp / 1.73 * (1.0 + pert * smooth_noise())
    vs[:water] = 0.0
    rho = 310.0 * vp ** 0.25 * (1.0 + pert * smooth_noise())
    return vp.astype(np.float32), vs.astype(np.float32), rho.astype(np.float32)
