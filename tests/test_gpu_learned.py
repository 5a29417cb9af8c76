"""The learned preconditioner end to end at a small size: training pairs
from the batched generator, AdamW training with the HIP kernels, and FWI
with the network direction and its plain-gradient fallback.

What is asserted: training lowers the loss on the pairs below that of the
zero direction (the network has learned part of true - start from the
gradient and model), and the preconditioned FWI never increases the misfit
(every accepted step reduces it; otherwise the plain step is taken).  The
gains over plain FWI at equal wall-clock are printed, and measured at full
size by the benchmark's preconditioner workload."""
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


def test_training_and_preconditioned_fwi_small():
    from seistorch_amd import fwi, generator, learned, precond
    spec = generator.GenSpec(nz=(64, 160), nx=(64, 192))
    pairs = learned.make_pairs(24, "cuda", spec=spec)
    n = len(pairs)
    assert n == 24
    p = precond.Params("cuda", seed=0)
    net = precond.UNet(p)
    opt = precond.AdamW(p, lr=1e-3)
    loss0, zero = learned.evaluate(net, pairs)
    learned.train(net, opt, pairs, 300, batch=8)
    loss1, zero = learned.evaluate(net, pairs)
    print(f"pairs {n}: loss {loss0:.4f} -> {loss1:.4f}, zero direction {zero:.4f}")
    assert loss1 < zero
    prob = _small_problem()
    _, hp = fwi.run_plain(prob, 6)
    _, hn = fwi.run_preconditioned(prob, net, 6)
    print("plain    misfit", [f"{m:.3e}" for m in hp.misfit], "rmse", [f"{r:.1f}" for r in hp.rmse])
    print("precond  misfit", [f"{m:.3e}" for m in hn.misfit], "rmse", [f"{r:.1f}" for r in hn.rmse],
          hn.kind)
    assert all(b <= a for a, b in zip(hn.misfit, hn.misfit[1:]))
    assert hn.misfit[-1] < hn.misfit[0]
