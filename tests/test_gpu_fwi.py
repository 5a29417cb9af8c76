"""The multiscale FWI loop (seistorch_dist.py:166-303 on one GPU): the
band-filtered L2 misfit's gradient through the HIP record filter and the
adjoint, and the Adam / steepest-descent updates per frequency scale.

The gradient through the filter is checked against FP64 autograd of the
reference restatement (finite differences of the FP32 misfit are too noisy
here: J+ - J- is ~1e-5 of J, where FP32 forward modelling varies by ~1e-6
of J between nearby models); the loop is checked to lower the filtered
misfit in every scale with either optimiser."""
import numpy as np
import pytest
import torch

from tests.test_gpu_learned import _small_problem

pytestmark = pytest.mark.gpu


def test_gradient_through_the_filter_matches_fp64_autograd():
    """d/dvp of <filtfilt(records), G>: the HIP filter's backward feeding the
    HIP adjoint, against FP64 autograd of the reference restatement with the
    record cotangent filtfilt(G) from the FP64 filter oracle (the filter is
    self-adjoint).  A fixed cotangent keeps FP32 record rounding out of the
    comparison (a misfit residual would amplify it)."""
    from oracle import ref_torch
    from oracle import signal as osg
    from seistorch_amd import signal as sg
    from seistorch_amd.equations2d import acoustic as ac
    from tests.cases import acoustic_case, rel_l2, to_dev
    c = acoustic_case(nshots=3, nt=300)
    t = to_dev(c)
    b, a = sg.butter(3, [4.0, 12.0], c.dt)
    vp = t["vp"].clone().requires_grad_(True)
    rec = ac.propagate(vp, t["d"], t["w"], t["src"], t["rec"], c.B, c.dt, c.h)
    G = np.random.default_rng(5).standard_normal(tuple(rec.shape)).astype(np.float32)
    (sg.filtfilt(rec, b, a) * torch.from_numpy(G).cuda()).sum().backward()
    g = vp.grad.cpu().numpy()
    cot = osg.filtfilt(b, a, G.astype(np.float64))
    _, g64, _ = ref_torch.gradient_ad(c.vp, c.d, c.dt, c.h, c.w, c.src, c.rec, None,
                                      dtype=torch.float64, rec_cotangent=cot)
    _, g32, _ = ref_torch.gradient_ad(c.vp, c.d, c.dt, c.h, c.w, c.src, c.rec, None,
                                      rec_cotangent=cot.astype(np.float32))
    I = (slice(50, -50), slice(50, -50))
    e64, e32, e_ref = (rel_l2(g[I], g64.numpy()[I]), rel_l2(g[I], g32.numpy()[I]),
                       rel_l2(g32.numpy()[I], g64.numpy()[I]))
    print(f"gradient through the filter: HIP vs FP64 AD {e64:.2e}, HIP vs FP32 AD {e32:.2e}, "
          f"FP32 AD vs FP64 AD {e_ref:.2e}")
    assert e32 < 1e-5 and e64 <= 2 * e_ref + 1e-6


@pytest.mark.parametrize("optimizer", ["adam", "steepestdescent"])
def test_multiscale_loop_lowers_each_band_misfit(optimizer):
    from seistorch_amd import fwi
    p = _small_problem()
    scales = [[2.0, 6.0], 10.0]
    vp, hist = fwi.run_multiscale(p, scales, epochs=4, lr=10.0, optimizer=optimizer,
                                  log=print)
    assert p.band == "all"  # restored
    for si, band in enumerate(scales):
        ms = [m for m, k in zip(hist.misfit, hist.kind) if k == f"scale{si}"]
        p.band = band
        final = p.misfit(vp) if si == len(scales) - 1 else None
        p.band = "all"
        print(f"{optimizer} band {band}: misfit {ms[0]:.4e} -> {ms[-1]:.4e}"
              + (f" (after the loop {final:.4e})" if final is not None else ""))
        assert ms[-1] < ms[0]


def test_multiscale_loop_with_smoothing_and_cut():
    """--grad-smooth, --grad-cut and --source-illumination in the loop: the
    misfit still drops, and the model stays at its start above the seabed (the
    cut zeroes the update there)."""
    import torch
    from seistorch_amd import fwi
    p = _small_problem()
    vi0 = p.interior(p.vp_init).clone()
    seabed = torch.ones_like(vi0)
    seabed[:8] = 0.0
    sm = {"counts": 2, "sigma": {"z": 1.0, "x": 1.0}, "radius": {"z": 2, "x": 2}}
    vp, hist = fwi.run_multiscale(p, ["all"], epochs=4, lr=10.0, smooth=sm, seabed=seabed,
                                  source_illumination=True, log=print)
    assert hist.misfit[-1] < hist.misfit[0]
    assert torch.equal(p.interior(vp)[:8], vi0[:8])


def test_habc_problem_in_the_loop():
    """equation: acoustic_habc in the FWI loop (its records, gradient and
    illumination), with smoothing and illumination preconditioning."""
    from seistorch_amd import fwi
    from seistorch_amd.habc import generate_habc_coefficients_2d
    p = _small_problem()
    p.equation = "acoustic_habc"
    p.d = generate_habc_coefficients_2d(tuple(p.vp_true.shape), 50).cuda()
    p.obs = p.records(p.vp_true)
    sm = {"counts": 1, "sigma": {"z": 1.0, "x": 1.0}, "radius": {"z": 2, "x": 2}}
    vp, hist = fwi.run_multiscale(p, ["all"], epochs=3, lr=10.0, smooth=sm,
                                  source_illumination=True, log=print)
    assert hist.misfit[-1] < hist.misfit[0]
