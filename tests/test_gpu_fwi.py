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


def test_run_from_config_end_to_end(tmp_path):
    """Files only: models, geometry and observed records written the
    reference's way, then the configuration-driven loop (two bands, smoothing,
    seabed cut, illumination) lowers the misfit of each band."""
    import numpy as np
    import torch
    from seistorch_amd import fwi, synthetic
    from seistorch_amd import io as sio
    from seistorch_amd.equations2d import acoustic as ac
    from seistorch_amd.pml import generate_pml_coefficients_2d
    from seistorch_amd.utils import pad_model, ricker_wave
    nz_i, nx_i, nt, dt, h = 80, 160, 600, 1e-3, 10.0
    true, init = synthetic.anomaly_model(nz_i, nx_i, seed=5, smooth=8.0)
    np.save(tmp_path / "true.npy", true)
    np.save(tmp_path / "init.npy", init)
    np.save(tmp_path / "seabed.npy", (np.arange(nz_i)[:, None] >= 6) * np.ones((1, nx_i)))
    sources = [[10 + 28 * i, 1] for i in range(5)]
    receivers = [[list(range(nx_i)), [2] * nx_i] for _ in range(5)]
    sio.write_pkl(tmp_path / "sources.pkl", sources)
    sio.write_pkl(tmp_path / "receivers.pkl", receivers)
    cfg = {"equation": "acoustic",
           "geom": {"truePath": {"vp": str(tmp_path / "true.npy")},
                    "initPath": {"vp": str(tmp_path / "init.npy")},
                    "sources": str(tmp_path / "sources.pkl"),
                    "receivers": str(tmp_path / "receivers.pkl"),
                    "seabed": str(tmp_path / "seabed.npy"), "obsPath": str(tmp_path / "obs.npy"),
                    "wavelet": None, "fm": 10.0, "dt": dt, "nt": nt, "h": h,
                    "wavelet_delay": 120, "wavelet_inverse": False, "multiple": False,
                    "boundary": {"type": "pml", "width": 50}, "multiscale": [[2.0, 6.0], "all"]},
           "training": {"N_epochs": 3, "lr": {"vp": 10.0}, "scale_decay": 0.85,
                        "lr_decay": 0.995, "filter_ord": 3,
                        "smooth": {"counts": 1, "sigma": {"z": 1.0, "x": 1.0},
                                   "radius": {"z": 2, "x": 2}}}}
    # forward modelling of the observed records, stored as the reference does
    vt = torch.from_numpy(pad_model(true).astype(np.float32)).cuda()
    nz, nx = vt.shape
    src, rec = sio.acquisition_from_cfg(cfg)
    obs = ac.propagate(vt, generate_pml_coefficients_2d((nz, nx), 50).cuda(),
                       ricker_wave(10.0, dt, nt, 120).cuda(), torch.from_numpy(src).cuda(),
                       torch.from_numpy(synthetic.rec_flat_index(rec, nz, nx)).cuda(), 5, dt, h)
    store = sio.SeisRecord(cfg)
    store.setup("forward")
    for i, shot in enumerate(sio.split_records(obs.cpu().numpy(), [len(r) for r in rec])):
        store[i] = shot
    store.write()
    p, vp, hist = fwi.run_from_config(cfg, grad_smooth=True, grad_cut=True,
                                      source_illumination=True, log=print)
    for si in range(2):
        ms = [m for m, k in zip(hist.misfit, hist.kind) if k == f"scale{si}"]
        assert ms[-1] < ms[0]
    # above the seabed (6 rows) the model keeps its starting values
    assert torch.equal(p.interior(vp)[:6], torch.from_numpy(init.astype(np.float32)).cuda()[:6])


@pytest.mark.parametrize("optimizer", ["adam", "steepestdescent"])
def test_elastic_multiscale_loop(optimizer):
    """Elastic FWI (vp and vs inverted, rho kept) through the HIP elastic
    forward and adjoint: the filtered misfit drops in each band."""
    import numpy as np
    import torch
    from scipy.ndimage import gaussian_filter
    from seistorch_amd import fwi
    from seistorch_amd.utils import pad_model
    from tests.cases import elastic_case, elastic_dev
    c = elastic_case(nshots=3, nt=400)
    t = elastic_dev(c)
    true = {"vp": t["vp"], "vs": t["vs"], "rho": t["rho"]}
    init = dict(true)
    for k in ("vp", "vs"):
        a = getattr(c, k)[50:-50, 50:-50].astype(np.float64)
        s = gaussian_filter(a, 4.0)
        if k == "vs":
            s[:6] = 0.0  # the water layer keeps vs = 0
        init[k] = torch.from_numpy(pad_model(s).astype(np.float32)).cuda()
    p = fwi.ElasticProblem(true, init, t["d"], t["w"], t["src"], t["rec"], c.B, c.dt, c.h)
    p.obs = p.records(true)
    models, hist = fwi.run_multiscale_elastic(p, [[5.0, 25.0], "all"], epochs=3,
                                              lr={"vp": 10.0, "vs": 6.0}, optimizer=optimizer,
                                              log=print)
    for si in range(2):
        ms = [m for m, k in zip(hist.misfit, hist.kind) if k == f"scale{si}"]
        assert ms[-1] < ms[0]
    assert torch.equal(models["rho"], true["rho"])


def test_run_from_config_elastic(tmp_path):
    """equation: elastic from files: vp, vs, rho models, geometry, two-component
    observed records; vp and vs inverted with their own learning rates."""
    import numpy as np
    import torch
    from scipy.ndimage import gaussian_filter
    from seistorch_amd import fwi, synthetic
    from seistorch_amd import io as sio
    from tests.cases import elastic_case
    c = elastic_case(nshots=3, nt=400)
    nz_i, nx_i = c.nz - 100, c.nx - 100
    paths = {}
    for k in ("vp", "vs", "rho"):
        a = getattr(c, k)[50:-50, 50:-50].astype(np.float64)
        np.save(tmp_path / f"true_{k}.npy", a)
        s = gaussian_filter(a, 4.0) if k != "rho" else a
        if k == "vs":
            s[:6] = 0.0
        np.save(tmp_path / f"init_{k}.npy", s)
        paths[k] = (str(tmp_path / f"true_{k}.npy"), str(tmp_path / f"init_{k}.npy"))
    # the case's acquisition in grid units of the unpadded model, [x, z]
    sources = [[int(x) - 50, int(z) - 50] for z, x in c.src]
    receivers = [[[int(x) - 50 for x in r[:, 1]], [int(z) - 50 for z in r[:, 0]]] for r in c.rec]
    sio.write_pkl(tmp_path / "s.pkl", sources)
    sio.write_pkl(tmp_path / "r.pkl", receivers)
    cfg = {"equation": "elastic",
           "geom": {"truePath": {k: v[0] for k, v in paths.items()},
                    "initPath": {k: v[1] for k, v in paths.items()},
                    "sources": str(tmp_path / "s.pkl"), "receivers": str(tmp_path / "r.pkl"),
                    "obsPath": str(tmp_path / "obs.npy"), "wavelet": None, "fm": 20.0,
                    "dt": c.dt, "nt": c.nt, "h": c.h, "wavelet_delay": 40,
                    "wavelet_inverse": False, "multiple": False,
                    "boundary": {"type": "pml", "width": 50}, "multiscale": ["all"],
                    "invlist": {"vp": True, "vs": True, "rho": False},
                    "source_type": ["txx", "tzz"], "receiver_type": ["vx", "vz"]},
           "training": {"N_epochs": 3, "lr": {"vp": 10.0, "vs": 6.0},
                        "smooth": {"counts": 1, "sigma": {"z": 1.0, "x": 1.0},
                                   "radius": {"z": 2, "x": 2}}}}
    np.save(tmp_path / "seabed.npy", (np.arange(nz_i)[:, None] >= 6) * np.ones((1, nx_i)))
    cfg["geom"]["seabed"] = str(tmp_path / "seabed.npy")
    # no records yet: model them from the true models
    p = fwi.problem_from_config({**cfg, "geom": {**cfg["geom"], "obsPath": None}})
    obs = p.records(p.models_true).cpu().numpy()
    store = sio.SeisRecord(cfg)
    store.setup("forward")
    for i, shot in enumerate(sio.split_records(obs, [len(r[0]) for r in receivers])):
        store[i] = shot
    store.write()
    p, models, hist = fwi.run_from_config(cfg, grad_smooth=True, grad_cut=True,
                                          source_illumination=True, log=print)
    assert hist.misfit[-1] < hist.misfit[0]
    assert torch.equal(models["rho"], p.models_init["rho"])
    # above the seabed (6 rows) vp and vs keep their starting values
    for k in ("vp", "vs"):
        assert torch.equal(p.interior(models[k])[:6], p.interior(p.models_init[k])[:6])


def test_run_from_config_variable_density(tmp_path):
    """equation: acoustic_rho_habc from files: vp and rho models, geometry,
    observed records; both inverted through the variable-density adjoint
    (RhoProblem), each with its own learning rate; the misfit falls and the
    rows above the seabed keep their starting values."""
    import numpy as np
    import torch
    from scipy.ndimage import gaussian_filter
    from seistorch_amd import fwi, synthetic
    from seistorch_amd import io as sio
    nz_i, nx_i, nt, dt, h = 60, 140, 450, 1e-3, 10.0
    vp, _ = synthetic.anomaly_model(nz_i, nx_i, vmin=1800.0, vmax=3000.0, seed=9, smooth=6.0)
    zz, xx = np.mgrid[0:nz_i, 0:nx_i]
    rho = 1800.0 + 4.0 * zz + 300.0 * (((zz - 35) / 8.0) ** 2 + ((xx - 70) / 15.0) ** 2 < 1.0)
    paths = {}
    for k, a in (("vp", vp.astype(np.float64)), ("rho", rho.astype(np.float64))):
        np.save(tmp_path / f"true_{k}.npy", a)
        np.save(tmp_path / f"init_{k}.npy", gaussian_filter(a, 5.0, mode="nearest"))
        paths[k] = (str(tmp_path / f"true_{k}.npy"), str(tmp_path / f"init_{k}.npy"))
    sources = [[15 + 55 * i, 1] for i in range(3)]
    receivers = [[list(range(nx_i)), [2] * nx_i] for _ in range(3)]
    sio.write_pkl(tmp_path / "s.pkl", sources)
    sio.write_pkl(tmp_path / "r.pkl", receivers)
    np.save(tmp_path / "seabed.npy", (np.arange(nz_i)[:, None] >= 6) * np.ones((1, nx_i)))
    cfg = {"equation": "acoustic_rho_habc",
           "geom": {"truePath": {k: v[0] for k, v in paths.items()},
                    "initPath": {k: v[1] for k, v in paths.items()},
                    "sources": str(tmp_path / "s.pkl"), "receivers": str(tmp_path / "r.pkl"),
                    "seabed": str(tmp_path / "seabed.npy"), "obsPath": str(tmp_path / "obs.npy"),
                    "wavelet": None, "fm": 15.0, "dt": dt, "nt": nt, "h": h,
                    "wavelet_delay": 80, "wavelet_inverse": False, "multiple": False,
                    "boundary": {"type": "habc", "width": 50}, "multiscale": ["all"],
                    "invlist": {"vp": True, "rho": True}},
           "training": {"N_epochs": 4, "lr": {"vp": 10.0, "rho": 10.0},
                        "smooth": {"counts": 1, "sigma": {"z": 1.0, "x": 1.0},
                                   "radius": {"z": 2, "x": 2}}}}
    p = fwi.problem_from_config({**cfg, "geom": {**cfg["geom"], "obsPath": None}})
    assert isinstance(p, fwi.RhoProblem) and p.invlist == ("vp", "rho")
    obs = p.records(p.models_true).cpu().numpy()
    store = sio.SeisRecord(cfg)
    store.setup("forward")
    for i, shot in enumerate(sio.split_records(obs, [len(r[0]) for r in receivers])):
        store[i] = shot
    store.write()
    p, models, hist = fwi.run_from_config(cfg, grad_smooth=True, grad_cut=True,
                                          source_illumination=True, log=print)
    print("variable-density FWI misfit", [f"{m:.3e}" for m in hist.misfit], flush=True)
    assert hist.misfit[-1] < hist.misfit[0]
    for k in ("vp", "rho"):
        assert not torch.equal(models[k], p.models_init[k])
        assert torch.equal(p.interior(models[k])[:6], p.interior(p.models_init[k])[:6])
