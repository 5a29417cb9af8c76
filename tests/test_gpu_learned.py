"""The learned preconditioner end to end at a small size: training pairs
from the batched generator, AdamW training with the HIP kernels, and FWI
with the network direction and its plain-gradient fallback.

What is asserted: on held-out pairs (generator models the network never saw)
the network's prediction of true - start is closer to it than the best
multiple of the gradient, i.e. than the update direction of the method
without the network with its step chosen in hindsight (squared error
relative to that of the zero direction); and the preconditioned FWI never
increases the misfit (every accepted step reduces it; otherwise the plain
step is taken).  The wall-clock comparison with plain FWI is measured at
full size by the benchmark's preconditioner workload."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _small_problem():
    import numpy as np
    from seistorch_amd import fwi, synthetic
    from seistorch_amd.pml import generate_pml_coefficients_2d
    from seistorch_amd.utils import pad_model, ricker_wave
    nz_i, nx_i, B, nt, dt, h, fm = 96, 192, 6, 700, 1e-3, 10.0, 10.0
    true, init = synthetic.anomaly_model(nz_i, nx_i, seed=7, smooth=8.0)
    src, rec = synthetic.surface_acquisition(nx_i, B, src_depth=2)
    vt, vi = pad_model(true), pad_model(init)
    nz, nx = vt.shape
    dev = "cuda"
    p = fwi.Problem(torch.from_numpy(vt).to(dev), torch.from_numpy(vi).to(dev),
                    generate_pml_coefficients_2d((nz, nx), 50).to(dev),
                    ricker_wave(fm, dt, nt, delay=120).to(dev),
                    torch.from_numpy(src.astype(np.int32)).to(dev),
                    torch.from_numpy(synthetic.rec_flat_index(rec, nz, nx)).to(dev), B, dt, h)
    p.obs = p.records(p.vp_true)
    return p


def test_network_gains_on_held_out_pairs_and_fwi_falls_back():
    from seistorch_amd import fwi, generator, learned, precond
    spec = generator.GenSpec(nz=(64, 160), nx=(64, 192))
    pairs = learned.make_pairs(96, "cuda", spec=spec)
    assert len(pairs) == 96
    val, trn = pairs[:16], pairs[16:]
    p = precond.Params("cuda", seed=0)
    net = precond.UNet(p)
    opt = precond.AdamW(p, lr=1e-3)
    learned.train(net, opt, trn, 600, batch=8)
    z, gs, nn, nb = learned.heldout_errors(net, val)
    print(f"held-out pairs {len(val)}: squared error / zero direction: best multiple of the "
          f"gradient {gs / z:.4f}, network {nn / z:.4f}, best multiple of the network "
          f"{nb / z:.4f}", flush=True)
    assert nn < 0.9 * gs and nn < z
    prob = _small_problem()
    _, hp = fwi.run_plain(prob, 6)
    _, hn = fwi.run_preconditioned(prob, net, 6)
    print("plain    misfit", [f"{m:.3e}" for m in hp.misfit], "rmse", [f"{r:.1f}" for r in hp.rmse])
    print("precond  misfit", [f"{m:.3e}" for m in hn.misfit], "rmse", [f"{r:.1f}" for r in hn.rmse],
          hn.kind)
    assert all(b <= a for a, b in zip(hn.misfit, hn.misfit[1:]))
    assert hn.misfit[-1] < hn.misfit[0]
