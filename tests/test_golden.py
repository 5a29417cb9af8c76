"""The oracle and the host-side generators against golden vectors of the
reference's own step functions (tests/golden/*.npz, written by
tests/golden/make_golden.py from the reference files).

These pin what closed forms cannot: the HABC strip/rotation/corner semantics
(acoustic_habc.py:79-221), the elastic staggering (elastic.py:7-37 with
utils.py:3-48), the variable-density face averaging
(acoustic_rho_habc.py:31-57), the ring layout (utils.py:50-89) and the HABC
weights and masks (habc.py:4-81).  FP64 oracles must agree to rounding level,
the FP32 restatements (oracle/ref_torch.py, which the HIP kernels equal bit
for bit) to the last bit.
"""
import os

import numpy as np
import pytest
import torch

from oracle import elastic as oel
from oracle import habc as ohabc
from oracle import ref_torch as rt
from oracle import coeffs

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(GOLD, f"{name}.npz"))


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# ------------------------------------------------------------------ HABC

@pytest.mark.parametrize("mult", [0, 1])
def test_habc_step_oracle_fp64(mult):
    g = load("habc_step")
    y = ohabc.step(g["vp"], g["h1"], g["h2"], g[f"b_m{mult}"], float(g["dt"]), float(g["h"]),
                   multiple=bool(mult))
    assert rel(y, g[f"y64_m{mult}"]) < 1e-13


@pytest.mark.parametrize("mult", [0, 1])
def test_habc_step_restatement_fp32(mult):
    g = load("habc_step")
    f = lambda k: torch.from_numpy(g[k])
    y, _ = rt.habc_step(f("vp"), f("h1"), f("h2"), torch.tensor(g["dt"]), torch.tensor(g["h"]),
                        f(f"b_m{mult}"), multiple=bool(mult))
    assert np.array_equal(y.numpy(), g[f"y32_m{mult}"])


@pytest.mark.parametrize("mult", [0, 1])
def test_habc_loop_oracle(mult):
    g = load("habc_loop")
    src = g[f"src_m{mult}"]
    rz = int(g[f"rec_z_m{mult}"])
    nx = g["vp"].shape[1]
    rec = [np.stack([np.full(nx, rz), np.arange(nx)], 1) for _ in range(len(src))]
    out, fields = ohabc.forward(g["vp"], g[f"b_m{mult}"] if f"b_m{mult}" in g else
                                load("habc_step")[f"b_m{mult}"], float(g["dt"]), float(g["h"]),
                                g["w"], src, rec, multiple=bool(mult), keep_fields=True)
    r = np.stack([o[..., 0] for o in out], 1)              # (nt, B, nx)
    assert rel(r, g[f"rec64_m{mult}"]) < 1e-12
    assert rel(fields[-1], g[f"last64_m{mult}"]) < 1e-7    # stored rounded to FP32
    # the FP32 restatement follows the reference's FP32 loop to rounding
    r32 = rt.habc_forward(g["vp"], load("habc_step")[f"b_m{mult}"], float(g["dt"]),
                          float(g["h"]), g["w"], src, rec, multiple=bool(mult))
    r32 = torch.stack([o[..., 0] for o in r32], 1).numpy()
    assert np.array_equal(r32, g[f"rec32_m{mult}"])


@pytest.mark.parametrize("mult", [0, 1])
def test_habc_generators(mult):
    """Product weights/masks (seistorch_amd.habc) and the oracle's, against
    habc.py:4-81 as the reference computes them."""
    from seistorch_amd import habc as phabc
    g = load("habc_step")
    nz, nx = g["vp"].shape
    b = phabc.generate_habc_coefficients_2d((nz, nx), 50, multiple=bool(mult)).numpy()
    assert np.array_equal(b, g[f"b_m{mult}"])
    tm, bm, lm, rm = phabc.bound_mask(nz, nx, 50, return_idx=True, multiple=bool(mult))
    assert np.array_equal(bm.numpy(), g[f"mask_bottom_m{mult}"].astype(bool))
    assert np.array_equal(lm.numpy(), g[f"mask_left_m{mult}"].astype(bool))
    assert np.array_equal(rm.numpy(), g[f"mask_right_m{mult}"].astype(bool))
    if mult:
        assert tm is None
    else:
        assert np.array_equal(tm.numpy(), g["mask_top_m0"].astype(bool))
        ot, ob, ol, orr = coeffs.habc_masks(nz, nx)
        assert np.array_equal(ot.astype(bool), g["mask_top_m0"].astype(bool))
        assert np.array_equal(ob.astype(bool), g["mask_bottom_m0"].astype(bool))
        assert np.array_equal(ol.astype(bool), g["mask_left_m0"].astype(bool))
        assert np.array_equal(orr.astype(bool), g["mask_right_m0"].astype(bool))
    ob = coeffs.habc_coefficients_2d(nz, nx, multiple=bool(mult))
    assert np.max(np.abs(ob - g[f"b_m{mult}"])) < 1e-7


# ------------------------------------------------------------------ variable density

@pytest.mark.parametrize("mult", [0, 1])
def test_rho_step(mult):
    g = load("rho")
    b = load("habc_step")[f"b_m{mult}"]
    f = lambda k, dt=torch.float64: torch.from_numpy(g[k]).to(dt)
    y64, _ = rt.rho_habc_step(f("vp"), f("rho"), f("h1"), f("h2"), float(g["dt"]), float(g["h"]),
                              torch.from_numpy(b).double(), multiple=bool(mult))
    assert rel(y64.numpy(), g[f"y64_m{mult}"]) < 1e-13
    f32 = lambda k: torch.from_numpy(g[k])
    y32, _ = rt.rho_habc_step(f32("vp"), f32("rho"), f32("h1"), f32("h2"), float(g["dt"]),
                              float(g["h"]), torch.from_numpy(b), multiple=bool(mult))
    assert np.array_equal(y32.numpy(), g[f"y32_m{mult}"])


def test_rho_loop():
    g = load("rho")
    b = load("habc_step")["b_m0"]
    src = g["loop_src"]
    rz = int(g["loop_rec_z"])
    nx = g["vp"].shape[1]
    rec = [np.stack([np.full(nx, rz), np.arange(nx)], 1) for _ in range(len(src))]
    r64 = rt.rho_habc_forward(g["vp"], g["rho"], b, float(g["dt"]), float(g["h"]), g["loop_w"],
                              src, rec, dtype=torch.float64)
    r64 = torch.stack([o[..., 0] for o in r64], 1).numpy()
    assert rel(r64, g["loop_rec64"]) < 1e-12
    r32 = rt.rho_habc_forward(g["vp"], g["rho"], b, float(g["dt"]), float(g["h"]), g["loop_w"],
                              src, rec)
    r32 = torch.stack([o[..., 0] for o in r32], 1).numpy()
    assert np.array_equal(r32, g["loop_rec32"])


# ------------------------------------------------------------------ elastic

NAMES = ("vx", "vz", "txx", "tzz", "txz")


def test_elastic_step_oracle_fp64():
    g = load("elastic")
    k = oel.coefficients(*(g[n].astype(np.float64) for n in ("vp", "vs", "rho", "d")),
                         float(g["dt"]), float(g["h"]))
    new, _ = oel.step(k, *(g[n].astype(np.float64) for n in NAMES))
    for n, a in zip(NAMES, new):
        assert rel(a, g[f"{n}64"]) < 1e-13, n


def test_elastic_step_restatement_fp32():
    g = load("elastic")
    f = lambda k: torch.from_numpy(g[k])
    new = rt.elastic_step(f("vp"), f("vs"), f("rho"), *(f(n) for n in NAMES),
                          torch.tensor(g["dt"]), torch.tensor(g["h"]), f("d"))
    for n, a in zip(NAMES, new):
        assert np.array_equal(a.numpy(), g[f"{n}32"]), n


def test_elastic_loop_oracle():
    g = load("elastic")
    src = g["loop_src"]
    rz = int(g["loop_rec_z"])
    nx = g["vp"].shape[1]
    rec = [np.stack([np.full(nx, rz), np.arange(nx)], 1) for _ in range(len(src))]
    out = oel.forward(g["vp"], g["vs"], g["rho"], g["d"], float(g["dt"]), float(g["h"]),
                      g["loop_w"], src, rec)
    r = np.stack(out, 1)                                   # (nt, B, nx, 2)
    assert rel(r, g["loop_rec64"]) < 1e-12
    r32 = rt.elastic_forward(g["vp"], g["vs"], g["rho"], g["d"], float(g["dt"]), float(g["h"]),
                             g["loop_w"], src, rec)
    r32 = torch.stack(r32, 1).numpy()
    assert np.array_equal(r32, g["loop_rec32"])


# ------------------------------------------------------------------ rings

@pytest.mark.parametrize("mult", [0, 1])
def test_ring_helpers(mult):
    """seistorch_amd's ring helpers against utils.py:50-89."""
    from seistorch_amd.equations2d import utils as pu
    g = load("reverse")
    ring = pu.save_boundaries(torch.from_numpy(g["ring_src"]))
    for name, r in zip(("top", "bottom", "left", "right"), ring):
        assert np.array_equal(r.numpy(), g[f"ring_{name}"]), name
    out = pu.restore_boundaries(torch.from_numpy(g["h1"]).clone(), ring, multiple=bool(mult))
    assert np.array_equal(out.numpy(), g[f"restored_m{mult}"])
