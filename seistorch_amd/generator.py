"""Seeded uneven multi-model generator: many small acoustic FWI problems solved
as one batch (records of the true model, then the L2-misfit gradient at a
smoothed start model), e.g. to make (gradient, true - start) training pairs.

Each instance (seeded from one integer):
* grid nz x nx uniform in [64, 512] x [64, 768] at dx = 10 m (nx on a 4-cell
  step so that every padded width takes 4-wide vector access), plus the
  50-cell boundary on every side;
* 3-8 dipping layers with velocities uniform in [1500, 4500] m/s (sorted with
  depth), 0-2 normal/reverse faults offsetting the layers, 1-3 Gaussian
  anomalies of +-5-10 %, clipped to [1500, 4500];
* start model: the true model Gaussian-smoothed (sigma = 20 cells, as the
  800x2400 FWI problem's start model);
* 4-16 shots evenly spaced at depth 2 cells, a receiver at every surface cell;
* Ricker 5-15 Hz (delay 1.2/f), dt = 1 ms, nt = ceil(1.2 * diagonal / 1500 m/s / dt).
The reference has no generator of its own; its per-model pieces are the
Ricker wavelet (seistorch/utils.py:235-249), edge padding
(seistorch/geom.py:213-240) and the PML profile (seistorch/pml.py:29-59).
"""
import math
from dataclasses import dataclass

import numpy as np
import torch
from scipy.ndimage import gaussian_filter

from .pml import generate_pml_coefficients_2d
from .synthetic import DEFAULT_SEED, rec_flat_index, surface_acquisition
from .utils import pad_model, ricker_wave

PAD = 50


@dataclass
class GenSpec:
    n_models: int = 256
    nz: tuple = (64, 512)
    nx: tuple = (64, 768)
    layers: tuple = (3, 8)
    vmin: float = 1500.0
    vmax: float = 4500.0
    faults: tuple = (0, 2)
    anomalies: tuple = (1, 3)
    shots: tuple = (4, 16)
    fpeak: tuple = (5.0, 15.0)
    dt: float = 1e-3
    dx: float = 10.0
    nt_factor: float = 1.2
    v_ref: float = 1500.0
    smooth: float = 20.0
    nx_step: int = 4


def random_model(rng, nz, nx, spec):
    """True velocity (nz, nx) float32: dipping layers, faults, anomalies."""
    L = int(rng.integers(spec.layers[0], spec.layers[1] + 1))
    vel = np.sort(rng.uniform(spec.vmin, spec.vmax, L))
    depth0 = np.sort(rng.uniform(0.08 * nz, 0.95 * nz, L - 1))
    slope = rng.uniform(-0.15, 0.15, L - 1)
    z = np.arange(nz, dtype=np.float64)[:, None]
    x = np.arange(nx, dtype=np.float64)[None, :]
    shift = np.zeros((nz, nx))
    for _ in range(int(rng.integers(spec.faults[0], spec.faults[1] + 1))):
        x0 = rng.uniform(0.2, 0.8) * nx
        tan = math.tan(math.radians(rng.uniform(-30.0, 30.0)))
        throw = rng.uniform(-0.1, 0.1) * nz
        shift = np.where(x > x0 + (z - nz / 2) * tan, shift + throw, shift)
    zz = z - shift
    layer = np.zeros((nz, nx), np.int64)
    for k in range(L - 1):
        layer += zz > depth0[k] + slope[k] * (x - nx / 2)
    v = vel[layer]
    zg, xg = np.mgrid[0:nz, 0:nx]
    for _ in range(int(rng.integers(spec.anomalies[0], spec.anomalies[1] + 1))):
        cz, cx = rng.uniform(0.2, 0.9) * nz, rng.uniform(0.1, 0.9) * nx
        sig = rng.uniform(0.03, 0.1) * min(nz, nx)
        amp = rng.choice([-1.0, 1.0]) * rng.uniform(0.05, 0.10)
        v = v * (1.0 + amp * np.exp(-((zg - cz) ** 2 + (xg - cx) ** 2) / (2 * sig * sig)))
    return np.clip(v, spec.vmin, spec.vmax).astype(np.float32)


def make_instance(seed, spec=GenSpec()):
    """One seeded problem as host arrays (padded grids, padded source and
    receiver indices)."""
    rng = np.random.default_rng(seed)
    nz = int(rng.integers(spec.nz[0], spec.nz[1] + 1))
    nx = int(rng.integers(spec.nx[0] // spec.nx_step, spec.nx[1] // spec.nx_step + 1)) * spec.nx_step
    true = random_model(rng, nz, nx, spec)
    start = gaussian_filter(true.astype(np.float64), spec.smooth, mode="nearest").astype(np.float32)
    B = int(rng.integers(spec.shots[0], spec.shots[1] + 1))
    fm = float(rng.uniform(*spec.fpeak))
    diag = math.hypot(nz * spec.dx, nx * spec.dx)
    nt = int(math.ceil(spec.nt_factor * diag / spec.v_ref / spec.dt))
    src, rec = surface_acquisition(nx, B, src_depth=2)
    vt, vs = pad_model(true), pad_model(start)
    NZ, NX = vt.shape
    return dict(seed=seed, nz_i=nz, nx_i=nx, nz=NZ, nx=NX, B=B, nt=nt, fm=fm, dt=spec.dt,
                h=spec.dx, vp_true=vt, vp_start=vs,
                wavelet=ricker_wave(fm, spec.dt, nt, delay=int(round(1.2 / fm / spec.dt)),
                                    dtype="numpy"),
                src=src.astype(np.int32), rec_flat=rec_flat_index(rec, NZ, NX))


def make_instances(seed=DEFAULT_SEED, spec=GenSpec(), first=0):
    """Instances first .. first + spec.n_models - 1 of the seeded stream (the
    i-th instance's seed depends on seed and i only)."""
    ss = np.random.SeedSequence(seed)
    kids = ss.spawn(first + spec.n_models)[first:]
    return [make_instance(int(s.generate_state(1)[0]), spec) for s in kids]


def cell_updates(insts):
    """Sum over instances of shots * padded cells * nt (one propagation)."""
    return sum(it["B"] * it["nz"] * it["nx"] * it["nt"] for it in insts)


def to_device(insts, device):
    """Device tensors of every instance (models, PML, wavelet, acquisition)."""
    out = []
    for it in insts:
        d = dict(it)
        for k in ("vp_true", "vp_start", "wavelet"):
            d[k] = torch.from_numpy(np.ascontiguousarray(it[k])).to(device)
        d["d"] = generate_pml_coefficients_2d((it["nz"], it["nx"]), PAD).to(device)
        d["src"] = torch.from_numpy(it["src"]).to(device)
        d["rec"] = torch.from_numpy(it["rec_flat"]).to(device)
        out.append(d)
    return out


def batch_instances(dev_insts, which):
    from .equations2d.acoustic_batch import BatchInstance
    return [BatchInstance(it[which], it["d"], it["wavelet"], it["src"], it["rec"], it["B"],
                          it["dt"], it["h"]) for it in dev_insts]


def observed(dev_insts, use_graph=True, **launch):
    """Records of every true model (no gradient state)."""
    from .equations2d.acoustic_batch import propagate_batch
    with torch.no_grad():
        return propagate_batch(batch_instances(dev_insts, "vp_true"), use_graph=use_graph,
                               **launch)


def misfit_gradients(dev_insts, obs, use_graph=True, snapshot_every=None, **launch):
    """(per-instance L2 misfit, per-instance dJ/dvp at the start model)."""
    from .equations2d.acoustic_batch import propagate_batch
    insts = batch_instances(dev_insts, "vp_start")
    for it in insts:
        it.vp = it.vp.detach().clone().requires_grad_(True)
    recs = propagate_batch(insts, use_graph=use_graph, snapshot_every=snapshot_every, **launch)
    losses = [torch.sum((r - o) ** 2) for r, o in zip(recs, obs)]
    torch.autograd.backward(losses)
    return torch.stack([l.detach() for l in losses]), [it.vp.grad for it in insts]
