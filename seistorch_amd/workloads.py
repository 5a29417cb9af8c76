"""The seeded benchmark problems, as host arrays on the padded grid.

* ``forward_modelling``  200 x 400 interior at dx = 10 m: 4 layers 1500-4000
  m/s with a Gaussian-smoothed (sigma 8 cells) +-5 % perturbation, PML 50,
  10 Hz Ricker, dt = 1 ms, nt = 1500, 8 shots at depth 2, a receiver at
  every surface cell.
* ``fwi_gradient``       800 x 2400 at dx = 10 m: smooth layered 1500-4500 m/s
  with 3 elliptical +-10 % anomalies; start model = true model smoothed with
  sigma 20 cells; PML 50, 8 Hz Ricker, dt = 0.8 ms, nt = 5000, 64 shots,
  2400 surface receivers.
* ``elastic``            600 x 1800 at dx = 5 m: vp 1800-4500, vs = vp/1.73
  with vs = 0 in a 20-cell water layer, Gardner density, each with a seeded
  smooth perturbation; start model without the perturbations; PML 50,
  10 Hz Ricker on txx and tzz, dt = 0.5 ms, nt = 6000, 32 shots, 1800
  receivers recording vx and vz.
* ``precond_fwi``       the learned preconditioner's FWI problem: the grid,
  shots, source, time axis and PML of ``fwi_gradient`` with a true model
  drawn from the uneven generator's family at that size (layers, faults and
  Gaussian anomalies, ``generator.random_model``), the start model smoothed
  with sigma 20 cells as the generator's start models are; ``model=
  "anomaly"`` takes ``fwi_gradient``'s model instead.
The uneven multi-model batch lives in ``generator``.

Padding follows seistorch/geom.py:213-240 (edge padding by 50 cells) and
seistorch/setup.py:386-483 (grid coordinates shifted by the pad); the
wavelet is seistorch/utils.py:235-249 with the delay in samples.
"""
import numpy as np

from . import synthetic
from .pml import generate_pml_coefficients_2d
from .utils import pad_model, ricker_wave

PAD = 50


def _acquisition(P, nx_i, nshots):
    src, rec = synthetic.surface_acquisition(nx_i, nshots, src_depth=2)
    P["src"], P["rec"] = src, rec
    P["rec_flat"] = synthetic.rec_flat_index(rec, P["nz"], P["nx"])
    P["d"] = generate_pml_coefficients_2d((P["nz"], P["nx"]), PAD).numpy()


def forward_modelling(seed=synthetic.DEFAULT_SEED, nt=None):
    nz_i, nx_i, nshots, dt, h, fm = 200, 400, 8, 1e-3, 10.0, 10.0
    nt = nt or 1500
    vp = pad_model(synthetic.layered_model(nz_i, nx_i, 1500.0, 4000.0, nlayers=4, sigma=8.0,
                                           pert=0.05, seed=seed)).astype(np.float32)
    P = dict(nz_i=nz_i, nx_i=nx_i, nz=vp.shape[0], nx=vp.shape[1], B=nshots, nt=nt, dt=dt,
             h=h, fm=fm, vp=vp)
    P["w"] = ricker_wave(fm, dt, nt, delay=256).numpy()
    _acquisition(P, nx_i, nshots)
    return P


def fwi_gradient(seed=synthetic.DEFAULT_SEED, nt=None):
    nz_i, nx_i, nshots, dt, h, fm = 800, 2400, 64, 0.8e-3, 10.0, 8.0
    nt = nt or 5000
    true, init = synthetic.anomaly_model(nz_i, nx_i, seed=seed)
    vt, vi = pad_model(true).astype(np.float32), pad_model(init).astype(np.float32)
    P = dict(nz_i=nz_i, nx_i=nx_i, nz=vi.shape[0], nx=vi.shape[1], B=nshots, nt=nt, dt=dt,
             h=h, fm=fm, vp_true=vt, vp_init=vi)
    P["w"] = ricker_wave(fm, dt, nt, delay=256).numpy()
    _acquisition(P, nx_i, nshots)
    return P


def elastic(seed=synthetic.DEFAULT_SEED, nt=None):
    nz_i, nx_i, nshots, dt, h, fm = 600, 1800, 32, 0.5e-3, 5.0, 10.0
    nt = nt or 6000
    P = dict(nz_i=nz_i, nx_i=nx_i, nz=nz_i + 2 * PAD, nx=nx_i + 2 * PAD, B=nshots, nt=nt,
             dt=dt, h=h, fm=fm)
    for tag, pert in (("true", 0.03), ("init", 0.0)):
        vp, vs, rho = synthetic.elastic_models(nz_i, nx_i, pert=pert, seed=seed)
        for k, a in (("vp", vp), ("vs", vs), ("rho", rho)):
            P[f"{k}_{tag}"] = pad_model(a).astype(np.float32)
    P["w"] = ricker_wave(fm, dt, nt, delay=160).numpy()
    _acquisition(P, nx_i, nshots)
    return P


def precond_fwi(seed=11, nt=None, model="generator"):
    if model == "anomaly":
        return fwi_gradient(nt=nt)
    if model != "generator":
        raise ValueError(f"model must be 'generator' or 'anomaly', got {model!r}")
    from scipy.ndimage import gaussian_filter
    from . import generator
    nz_i, nx_i, nshots, dt, h, fm = 800, 2400, 64, 0.8e-3, 10.0, 8.0
    nt = nt or 5000
    spec = generator.GenSpec()
    true = generator.random_model(np.random.default_rng(seed), nz_i, nx_i, spec)
    init = gaussian_filter(true.astype(np.float64), spec.smooth, mode="nearest")
    vt, vi = pad_model(true).astype(np.float32), pad_model(init.astype(np.float32))
    P = dict(nz_i=nz_i, nx_i=nx_i, nz=vi.shape[0], nx=vi.shape[1], B=nshots, nt=nt, dt=dt,
             h=h, fm=fm, vp_true=vt, vp_init=vi.astype(np.float32))
    P["w"] = ricker_wave(fm, dt, nt, delay=256).numpy()
    _acquisition(P, nx_i, nshots)
    return P

