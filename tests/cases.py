"""Seeded small problems shared by the parity tests (sizes the oracle finishes in seconds)."""
import numpy as np
import torch

from seistorch_amd import synthetic
from seistorch_amd.pml import generate_pml_coefficients_2d
from seistorch_amd.utils import pad_model, ricker_wave


class Case:
    pass


def acoustic_case(nz_i=60, nx_i=90, nshots=3, nt=400, dt=1e-3, h=10.0, fm=15.0, seed=1,
                  vmin=1500.0, vmax=3000.0, delay=60, src_depth=2):
    c = Case()
    v = synthetic.layered_model(nz_i, nx_i, vmin, vmax, seed=seed)
    c.vp = pad_model(v).astype(np.float32)
    c.nz, c.nx = c.vp.shape
    c.d = generate_pml_coefficients_2d((c.nz, c.nx), 50).numpy()
    c.w = ricker_wave(fm, dt, nt, delay=delay).numpy()
    c.src, c.rec = synthetic.surface_acquisition(nx_i, nshots, src_depth=src_depth)
    c.dt, c.h, c.nt, c.B = dt, h, nt, nshots
    rng = np.random.default_rng(seed + 100)
    vt = v * (1.0 + 0.03 * (rng.random(v.shape) > 0.5) * (np.arange(nz_i)[:, None] > nz_i // 3))
    c.vp_true = pad_model(vt).astype(np.float32)
    return c


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def to_dev(c, dev="cuda"):
    t = {}
    t["vp"] = torch.from_numpy(c.vp).to(dev)
    t["d"] = torch.from_numpy(c.d).to(dev)
    t["w"] = torch.from_numpy(c.w).to(dev)
    t["src"] = torch.from_numpy(c.src.astype(np.int32)).to(dev)
    t["rec"] = torch.from_numpy(synthetic.rec_flat_index(c.rec, c.nz, c.nx)).to(dev)
    return t


def elastic_case(nz_i=40, nx_i=60, nshots=3, nt=300, dt=5e-4, h=5.0, fm=20.0, seed=2,
                 water=6, delay=40):
    """Layered vp 1800..3000, vs = vp/1.73 with a water layer (vs = 0) on top,
    Gardner density 310*vp^0.25, PML 50; explosive sources at depth 2."""
    c = Case()
    v = synthetic.layered_model(nz_i, nx_i, 1800.0, 3000.0, nlayers=4, seed=seed)
    vs = v / 1.73
    vs[:water] = 0.0
    rho = 310.0 * v ** 0.25
    c.vp = pad_model(v).astype(np.float32)
    c.vs = pad_model(vs).astype(np.float32)
    c.rho = pad_model(rho).astype(np.float32)
    c.nz, c.nx = c.vp.shape
    c.d = generate_pml_coefficients_2d((c.nz, c.nx), 50).numpy()
    c.w = ricker_wave(fm, dt, nt, delay=delay).numpy()
    c.src, c.rec = synthetic.surface_acquisition(nx_i, nshots, src_depth=2)
    c.dt, c.h, c.nt, c.B = dt, h, nt, nshots
    return c


def elastic_dev(c, dev="cuda"):
    t = to_dev(c, dev)
    t["vs"] = torch.from_numpy(c.vs).to(dev)
    t["rho"] = torch.from_numpy(c.rho).to(dev)
    return t
