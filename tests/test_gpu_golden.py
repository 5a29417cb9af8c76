"""The HIP drop-in functions on the inputs of the golden vectors, against the
reference's own outputs (tests/golden/*.npz, written from the reference's step
functions by tests/golden/make_golden.py).

Bar: bit-identical FP32 results for every step, reverse step and record, since
the kernels round op by op in the reference's order.  FP64 errors of the FP32
records are printed for the record.
"""
import os

import numpy as np
import pytest
import torch

from seistorch_amd import synthetic
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
NAMES = ("vx", "vz", "txx", "tzz", "txz")


def load(name):
    return np.load(os.path.join(GOLD, f"{name}.npz"))


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _masks(nz, nx, B, mult):
    from seistorch_amd.habc import bound_mask
    return bound_mask(nz, nx, 50, "cuda", B, return_idx=True, multiple=bool(mult))


def _same(name, got, want):
    got = got.detach().cpu().numpy() if torch.is_tensor(got) else np.asarray(got)
    nbad = int(np.sum(got != want))
    print(f"{name}: {nbad} of {want.size} values differ, rel-L2 {rel_l2(got, want):.2e}",
          flush=True)
    assert nbad == 0, name


@pytest.mark.parametrize("mult", [0, 1])
def test_habc_step(mult):
    from seistorch_amd.equations2d import acoustic_habc as ah
    g = load("habc_step")
    B, nz, nx = g["h1"].shape
    y, h1 = ah._time_step(dev(g["vp"]), dev(g["h1"]), dev(g["h2"]), torch.tensor(g["dt"]),
                          torch.tensor(g["h"]), dev(g[f"b_m{mult}"]),
                          habcs=_masks(nz, nx, B, mult))
    _same(f"habc step multiple={mult}", y, g[f"y32_m{mult}"])


@pytest.mark.parametrize("mult", [0, 1])
def test_rho_step(mult):
    from seistorch_amd.equations2d import acoustic_rho_habc as ar
    g = load("rho")
    b = load("habc_step")[f"b_m{mult}"]
    B, nz, nx = g["h1"].shape
    y, _ = ar._time_step(dev(g["vp"]), dev(g["rho"]), dev(g["h1"]), dev(g["h2"]),
                         torch.tensor(g["dt"]), torch.tensor(g["h"]), dev(b),
                         habcs=_masks(nz, nx, B, mult))
    _same(f"rho step multiple={mult}", y, g[f"y32_m{mult}"])


def test_elastic_step():
    from seistorch_amd.equations2d import elastic as el
    g = load("elastic")
    out = el._time_step(dev(g["vp"]), dev(g["vs"]), dev(g["rho"]), *(dev(g[n]) for n in NAMES),
                        torch.tensor(g["dt"]), torch.tensor(g["h"]), dev(g["d"]))
    for n, a in zip(NAMES, out):
        _same(f"elastic step {n}", a, g[f"{n}32"])


def _src(smask):
    def f(Y, X, dt=1.0):
        Y_new = Y.clone()
        Y_new += smask * X
        return Y_new
    return f


def _rings(g):
    from seistorch_amd.equations2d.utils import save_boundaries
    rs = dev(g["ring_src"])
    return [save_boundaries(rs), save_boundaries(rs)]


@pytest.mark.parametrize("mult", [0, 1])
def test_acoustic_reverse_step(mult):
    from seistorch_amd.equations2d import acoustic as ac
    g = load("reverse")
    fn = ac._time_step_backward_multiple if mult else ac._time_step_backward
    y, _ = fn(dev(g["vp"]), dev(g["h1"]), dev(g["h2"]), torch.tensor(g["dt"]),
              torch.tensor(g["h"]), dev(g[f"d_m{mult}"]), _rings(g),
              [["h1"], _src(dev(g["smask"])), torch.tensor(g["wv"]).cuda()])
    _same(f"acoustic reverse step multiple={mult}", y,
          g["acoustic_back_multiple" if mult else "acoustic_back"])


@pytest.mark.parametrize("mult", [0, 1])
def test_habc_reverse_step(mult):
    from seistorch_amd.equations2d import acoustic_habc as ah
    g = load("reverse")
    b = load("habc_step")[f"b_m{mult}"]
    fn = ah._time_step_backward_multiple if mult else ah._time_step_backward
    y, _ = fn(dev(g["vp"]), dev(g["h1"]), dev(g["h2"]), torch.tensor(g["dt"]),
              torch.tensor(g["h"]), dev(b), _rings(g),
              [["h1"], _src(dev(g["smask"])), torch.tensor(g["wv"]).cuda()])
    _same(f"habc reverse step multiple={mult}", y,
          g["habc_back_multiple" if mult else "habc_back"])


@pytest.mark.parametrize("mult", [0, 1])
def test_rho_reverse_step(mult):
    from seistorch_amd.equations2d import acoustic_rho_habc as ar
    g = load("reverse")
    b = load("habc_step")[f"b_m{mult}"]
    fn = ar._time_step_backward_multiple if mult else ar._time_step_backward
    y, _ = fn(dev(g["vp"]), dev(g["rho"]), dev(g["h1"]), dev(g["h2"]), torch.tensor(g["dt"]),
              torch.tensor(g["h"]), dev(b), _rings(g),
              [["h1"], _src(dev(g["smask"])), torch.tensor(g["wv"]).cuda()])
    _same(f"rho reverse step multiple={mult}", y, g["rho_back_multiple" if mult else "rho_back"])


@pytest.mark.parametrize("mult", [0, 1])
def test_elastic_reverse_step(mult):
    from seistorch_amd.equations2d import elastic as el
    from seistorch_amd.equations2d.utils import save_boundaries
    g = load("reverse")
    rings = [save_boundaries(dev(g[f"el_ring_{n}"])) for n in NAMES]
    smask = dev(g["smask"][:1])

    def src(Y, X, dt=1.0):
        Y += smask * X                                   # first-order equations: in place
        return Y
    fn = el._time_step_backward_multiple if mult else el._time_step_backward
    out = fn(dev(g["el_vp"]), dev(g["el_vs"]), dev(g["el_rho"]),
             *(dev(g[f"el_{n}"]) for n in NAMES), torch.tensor(g["el_dt"]),
             torch.tensor(g["el_h"]), dev(g[f"d_m{mult}"]), rings,
             [["txx", "tzz"], src, torch.tensor(g["el_wv"]).cuda()])
    for n, a in zip(NAMES, out):
        _same(f"elastic reverse step multiple={mult} {n}", a, g[f"el_back_m{mult}_{n}"])


def _row_receivers(B, rz, nz, nx):
    rec = [np.stack([np.full(nx, rz), np.arange(nx)], 1) for _ in range(B)]
    return dev(synthetic.rec_flat_index(rec, nz, nx))


@pytest.mark.parametrize("mult", [0, 1])
def test_habc_records(mult):
    from seistorch_amd.equations2d import acoustic_habc as ah
    g = load("habc_loop")
    b = load("habc_step")[f"b_m{mult}"]
    src = g[f"src_m{mult}"]
    B = len(src)
    nz, nx = g["vp"].shape
    r = ah.propagate(dev(g["vp"]), dev(b), dev(g["w"]), dev(src.astype(np.int32)),
                     _row_receivers(B, int(g[f"rec_z_m{mult}"]), nz, nx), B, float(g["dt"]),
                     float(g["h"]), bool(mult))
    got = r.reshape(r.shape[0], B, nx)
    print(f"habc records multiple={mult}: vs reference FP64 "
          f"{rel_l2(got.cpu().numpy(), g[f'rec64_m{mult}']):.2e}")
    _same(f"habc records multiple={mult}", got, g[f"rec32_m{mult}"])


def test_rho_records():
    from seistorch_amd.equations2d import acoustic_rho_habc as ar
    g = load("rho")
    b = load("habc_step")["b_m0"]
    src = g["loop_src"]
    B = len(src)
    nz, nx = g["vp"].shape
    r = ar.propagate(dev(g["vp"]), dev(g["rho"]), dev(b), dev(g["loop_w"]),
                     dev(src.astype(np.int32)), _row_receivers(B, int(g["loop_rec_z"]), nz, nx),
                     B, float(np.float32(g["dt"])), float(np.float32(g["h"])))
    got = r.reshape(r.shape[0], B, nx)
    print(f"rho records: vs reference FP64 {rel_l2(got.cpu().numpy(), g['loop_rec64']):.2e}")
    _same("rho records", got, g["loop_rec32"])


def test_elastic_records():
    from seistorch_amd.equations2d import elastic as el
    g = load("elastic")
    src = g["loop_src"]
    B = len(src)
    nz, nx = g["vp"].shape
    r = el.propagate(dev(g["vp"]), dev(g["vs"]), dev(g["rho"]), dev(g["d"]), dev(g["loop_w"]),
                     dev(src.astype(np.int32)), _row_receivers(B, int(g["loop_rec_z"]), nz, nx),
                     B, float(g["dt"]), float(g["h"]))
    got = r.reshape(r.shape[0], B, nx, 2)
    print(f"elastic records: vs reference FP64 {rel_l2(got.cpu().numpy(), g['loop_rec64']):.2e}")
    _same("elastic records", got, g["loop_rec32"])
