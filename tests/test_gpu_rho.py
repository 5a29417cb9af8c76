"""Parity of the HIP variable-density acoustic path (acoustic_rho_habc.py)
with the FP32 restatement of the reference (oracle/ref_torch.rho_habc_*),
computed in the reference's operation order, and with its FP64 evaluation."""
import numpy as np
import pytest
import torch

from oracle import ref_torch
from tests.cases import acoustic_case, rel_l2, to_dev

pytestmark = pytest.mark.gpu


def _rho_case(nshots=3, nt=300, multiple=False, seed=1):
    from seistorch_amd.habc import generate_habc_coefficients_2d
    c = acoustic_case(nz_i=60, nx_i=80, nshots=nshots, nt=nt, seed=seed)
    if multiple:  # no top padding: drop the 50 top rows of the padded model
        c.vp = np.ascontiguousarray(c.vp[50:])
        c.nz = c.vp.shape[0]
        c.src = c.src.copy()
        c.src[:, 0] -= 50
        c.rec = [r - np.array([50, 0]) for r in c.rec]
    rng = np.random.default_rng(seed + 7)
    # density: layers with a rough jump, 1000-2600 kg/m^3
    z = np.arange(c.nz)[:, None] * np.ones((1, c.nx))
    c.rho = (1000.0 + 8.0 * z + 300.0 * (rng.random((c.nz, c.nx)) > 0.7)).astype(np.float32)
    c.b = generate_habc_coefficients_2d((c.nz, c.nx), 50, multiple=multiple).numpy()
    c.multiple = multiple
    return c


def _masks(c):
    from seistorch_amd.habc import bound_mask
    return bound_mask(c.nz, c.nx, 50, "cuda", c.B, return_idx=True, multiple=c.multiple)


@pytest.mark.parametrize("multiple", [False, True])
@pytest.mark.parametrize("B", [2, 5])
def test_step_matches_reference(multiple, B):
    from seistorch_amd.equations2d import acoustic_rho_habc as ar
    c = _rho_case(nshots=B, multiple=multiple)
    rng = np.random.default_rng(B)
    h1 = rng.standard_normal((B, c.nz, c.nx)).astype(np.float32)
    h2 = rng.standard_normal((B, c.nz, c.nx)).astype(np.float32)
    d = lambda a: torch.from_numpy(a).cuda()
    y, _ = ar._time_step(d(c.vp), d(c.rho), d(h1), d(h2), c.dt, c.h, d(c.b), habcs=_masks(c))
    t = lambda a, dt=torch.float32: torch.from_numpy(a).to(dt)
    yr, _ = ref_torch.rho_habc_step(t(c.vp), t(c.rho), t(h1), t(h2), c.dt, c.h, t(c.b),
                                    multiple=multiple)
    y64, _ = ref_torch.rho_habc_step(*(t(a, torch.float64) for a in (c.vp, c.rho, h1, h2)), c.dt,
                                     c.h, t(c.b, torch.float64), multiple=multiple)
    yc = y.cpu()
    bitwise = torch.equal(yc, yr)
    e32 = (yc - yr).abs().max().item() / yr.abs().max().item()
    e64 = rel_l2(yc.numpy(), y64.numpy())
    print(f"rho step multiple={multiple} B={B}: vs FP32 reference order max {e32:.2e} "
          f"(bitwise {bitwise}), vs FP64 {e64:.2e}")
    assert e32 < 1e-6 and e64 < 1e-6


def test_reverse_step_is_the_unblended_update():
    """_time_step_backward: the update without the strips, then the ring
    restore and the source (acoustic_rho_habc.py:59-89)."""
    from seistorch_amd.equations2d import acoustic_rho_habc as ar
    from seistorch_amd.equations2d.utils import save_boundaries
    c = _rho_case(nshots=2)
    rng = np.random.default_rng(3)
    h1, h2, prev = (rng.standard_normal((2, c.nz, c.nx)).astype(np.float32) for _ in range(3))
    d = lambda a: torch.from_numpy(a).cuda()
    ring = save_boundaries(d(prev))

    def src_func(y, v, sign):
        y = y.clone()
        y[:, 60, 70] += v
        return y

    y, _ = ar._time_step_backward(d(c.vp), d(c.rho), d(h1), d(h2), c.dt, c.h, d(c.b),
                                  (ring, None), [["h1"], src_func, 0.25])
    t = lambda a: torch.from_numpy(a.astype(np.float64))
    C = c.dt ** 2 * c.h ** (-2)
    ref = 2 * t(h1) - t(h2) + C * t(c.rho) * t(c.vp) ** 2 * ref_torch.rho_operator(t(c.rho), t(h1))
    ref[:, 60, 70] += 0.25
    inner = (slice(None), slice(50, -50), slice(50, -50))
    e = rel_l2(y.cpu().numpy()[inner], ref.numpy()[inner])
    ring_ok = torch.equal(y[:, 49, :], d(prev)[:, 49, :]) and \
        torch.equal(y[:, :, -50], d(prev)[:, :, -50])
    print(f"rho reverse step: interior vs FP64 {e:.2e}, ring restored {ring_ok}")
    assert e < 1e-6 and ring_ok


@pytest.mark.parametrize("multiple", [False, True])
def test_fused_forward_records(multiple):
    from seistorch_amd import synthetic
    from seistorch_amd.equations2d import acoustic_rho_habc as ar
    c = _rho_case(nshots=3, nt=400, multiple=multiple)
    t = to_dev(c)
    t["rec"] = torch.from_numpy(synthetic.rec_flat_index(c.rec, c.nz, c.nx)).cuda()
    rec = ar.propagate(t["vp"], torch.from_numpy(c.rho).cuda(), torch.from_numpy(c.b).cuda(),
                       t["w"], t["src"], t["rec"], c.B, c.dt, c.h, multiple=multiple)
    r32 = torch.cat(ref_torch.rho_habc_forward(c.vp, c.rho, c.b, c.dt, c.h, c.w, c.src, c.rec,
                                               multiple=multiple), 1)[..., 0]
    r64 = torch.cat(ref_torch.rho_habc_forward(c.vp, c.rho, c.b, c.dt, c.h, c.w, c.src, c.rec,
                                               dtype=torch.float64, multiple=multiple), 1)[..., 0]
    got = rec.cpu()
    e32, e64 = rel_l2(got.numpy(), r32.numpy()), rel_l2(got.numpy(), r64.numpy())
    print(f"rho records multiple={multiple}: vs FP32 reference order {e32:.2e} "
          f"(bitwise {torch.equal(got, r32)}), vs FP64 {e64:.2e}")
    assert e32 < 1e-5 and e64 < 1e-5


def test_constant_density_matches_the_acoustic_habc_records():
    from seistorch_amd.equations2d import acoustic_habc as ah
    from seistorch_amd.equations2d import acoustic_rho_habc as ar
    c = _rho_case(nshots=3, nt=400)
    t = to_dev(c)
    b = torch.from_numpy(c.b).cuda()
    rho = torch.full_like(t["vp"], 1800.0)
    r_rho = ar.propagate(t["vp"], rho, b, t["w"], t["src"], t["rec"], c.B, c.dt, c.h)
    r_ac = ah.propagate(t["vp"], b, t["w"], t["src"], t["rec"], c.B, c.dt, c.h)
    e = rel_l2(r_rho.cpu().numpy(), r_ac.cpu().numpy())
    print(f"constant density vs acoustic HABC records: {e:.2e}")
    # two FP32 formulations of one exact update (equal in FP64 to 1e-13, see
    # tests/test_oracle_rho.py): each carries ~5e-6 against FP64 after 400
    # steps, so they differ by up to the sum
    assert e < 2e-5


def _rho_grads(c, G, K=None):
    from seistorch_amd.equations2d import acoustic_rho_habc as ar
    t = to_dev(c)
    vp = t["vp"].clone().requires_grad_(True)
    rho = torch.from_numpy(c.rho).cuda().requires_grad_(True)
    rec = ar.propagate(vp, rho, torch.from_numpy(c.b).cuda(), t["w"], t["src"], t["rec"], c.B,
                       c.dt, c.h, multiple=c.multiple, snapshot_every=K)
    rec.backward(torch.from_numpy(G).cuda())
    return rec.detach().cpu().numpy(), vp.grad.cpu().numpy(), rho.grad.cpu().numpy()


@pytest.mark.parametrize("multiple,K", [(False, None), (False, 0), (True, None)])
def test_gradient_matches_autograd(multiple, K):
    """dJ/dvp and dJ/drho of J = sum(G * records) from the fused adjoint
    against autograd through the FP32 restatement (1e-5 bar with snapshots)
    and FP64 autograd, on the undamped interior (rows from 0 with a free
    surface); exactly zero elsewhere.  Ring-only (K = 0): bounded by the
    FP32 boundary-saving drift, as the reference's own path."""
    c = _rho_case(nshots=3, nt=300, multiple=multiple, seed=3)
    syn = torch.cat(ref_torch.rho_habc_forward(c.vp, c.rho, c.b, c.dt, c.h, c.w, c.src, c.rec,
                                               multiple=multiple), 1)[..., 0].numpy()
    G = (np.random.default_rng(5).standard_normal(syn.shape) * np.abs(syn).max()).astype(np.float32)
    rec, gvp, grho = _rho_grads(c, G, K)
    assert np.array_equal(rec, syn)
    (a32, r32), _ = ref_torch.rho_habc_gradient_ad(c.vp, c.rho, c.b, c.dt, c.h, c.w, c.src, c.rec,
                                                   G, multiple=multiple)
    (a64, r64), _ = ref_torch.rho_habc_gradient_ad(c.vp, c.rho, c.b, c.dt, c.h, c.w, c.src, c.rec,
                                                   G, dtype=torch.float64, multiple=multiple)
    region = (slice(0 if multiple else 50, -50), slice(50, -50))
    mask = np.ones_like(gvp, bool)
    mask[region] = False
    for name, g, g32, g64 in (("vp", gvp, a32.numpy(), a64.numpy()),
                              ("rho", grho, r32.numpy(), r64.numpy())):
        e_ref = rel_l2(g[region], g32[region])
        e_64 = rel_l2(g[region], g64[region])
        e_r64 = rel_l2(g32[region], g64[region])
        print(f"rho gradient d{name} multiple={multiple} K={K}: vs FP32 autograd {e_ref:.2e}, vs "
              f"FP64 {e_64:.2e}, FP32 autograd vs FP64 {e_r64:.2e}", flush=True)
        assert np.all(g[mask] == 0.0)
        if K == 0:
            assert e_64 < 20 * max(e_r64, 1e-6)
        else:
            assert e_ref < 1e-5 and e_64 <= 2 * e_r64 + 1e-6


def test_gradient_finite_differences():
    """FP32 central differences of J = sum(G * records) at interior cells for
    vp and rho."""
    from seistorch_amd.equations2d import acoustic_rho_habc as ar
    c = _rho_case(nshots=2, nt=300, seed=8)
    t = to_dev(c)
    rng = np.random.default_rng(11)
    R = t["rec"].numel()
    G = torch.from_numpy(rng.standard_normal((c.nt, R)).astype(np.float32)).cuda()
    _, gvp, grho = _rho_grads(c, G.cpu().numpy())
    b = torch.from_numpy(c.b).cuda()
    rho0 = torch.from_numpy(c.rho).cuda()
    for name, g, step in (("vp", gvp, 20.0), ("rho", grho, 20.0)):
        for i, j in ((60, 80), (75, 100), (95, 120)):
            vals = []
            for sgn in (1.0, -1.0):
                vp = t["vp"].clone()
                rho = rho0.clone()
                (vp if name == "vp" else rho)[i, j] += sgn * step
                r = ar.propagate(vp, rho, b, t["w"], t["src"], t["rec"], c.B, c.dt, c.h)
                vals.append(float((r.double() * G.double()).sum()))
            fd = (vals[0] - vals[1]) / (2 * step)
            print(f"rho FD d{name} cell ({i},{j}): fd {fd:.6e} adjoint {g[i, j]:.6e}", flush=True)
            assert abs(fd - g[i, j]) <= 1e-2 * max(abs(fd), 1e-2 * np.abs(g).max())


def test_illumination_matches_the_reference_loop():
    """The variable-density forward's source illumination against the
    reference loop on the per-step function (rnn.py:204-205 order)."""
    from seistorch_amd.equations2d import acoustic_rho_habc as ar
    c = _rho_case(nshots=3, nt=120)
    t = to_dev(c)
    d = lambda a: torch.from_numpy(a).cuda()
    illum = torch.empty((c.nz, c.nx), device="cuda")
    rec = ar.propagate(t["vp"], d(c.rho), d(c.b), t["w"], t["src"], t["rec"], c.B, c.dt, c.h,
                       illum=illum)
    rec0 = ar.propagate(t["vp"], d(c.rho), d(c.b), t["w"], t["src"], t["rec"], c.B, c.dt, c.h)
    assert torch.equal(rec, rec0)
    smask = torch.zeros((c.B, c.nz, c.nx), device="cuda")
    smask[torch.arange(c.B), t["src"][:, 0].long(), t["src"][:, 1].long()] = 1.0
    h1 = torch.zeros_like(smask)
    h2 = torch.zeros_like(smask)
    P = torch.zeros((c.nz, c.nx), device="cuda")
    for n in range(c.nt):
        h1, h2 = ar._time_step(t["vp"], d(c.rho), h1, h2, c.dt, c.h, d(c.b), habcs=_masks(c))
        h1 = h1 + smask * t["w"][n]
        s = torch.zeros_like(P)
        for k in range(c.B):
            s = s + h1[k] * h1[k]
        P = P + s
    e = rel_l2(illum.cpu().numpy(), P.cpu().numpy())
    print(f"rho illumination vs the per-step loop: {e:.2e} (bitwise {torch.equal(illum, P)})")
    assert e < 1e-6


def test_full_size_properties():
    """900 x 2500 padded grid, 64 shots, 150 steps: records are linear in the
    wavelet (scaling by 2 is exact in FP32 away from underflow), and a shot's
    records do not depend on the other shots of the batch (bit for bit)."""
    from seistorch_amd import synthetic
    from seistorch_amd.equations2d import acoustic_rho_habc as ar
    from seistorch_amd.habc import generate_habc_coefficients_2d
    from seistorch_amd.utils import pad_model, ricker_wave
    nz_i, nx_i, B, nt, dt, h = 800, 2400, 64, 150, 0.8e-3, 10.0
    true, _ = synthetic.anomaly_model(nz_i, nx_i, seed=11)
    vp = torch.from_numpy(pad_model(true).astype(np.float32)).cuda()
    nz, nx = vp.shape
    z = torch.arange(nz, device="cuda", dtype=torch.float32)[:, None]
    rho = (1000.0 + 1.2 * z).expand(nz, nx).contiguous()
    b = generate_habc_coefficients_2d((nz, nx), 50).cuda()
    src, rec = synthetic.surface_acquisition(nx_i, B, src_depth=2)
    w = ricker_wave(8.0, dt, nt, delay=60).cuda()
    s = torch.from_numpy(src.astype(np.int32)).cuda()
    flat = torch.from_numpy(synthetic.rec_flat_index(rec, nz, nx)).cuda()
    r1 = ar.propagate(vp, rho, b, w, s, flat, B, dt, h)
    r2 = ar.propagate(vp, rho, b, 2.0 * w, s, flat, B, dt, h)
    e_lin = rel_l2(r2.cpu().numpy(), 2.0 * r1.cpu().numpy())
    k = 37  # one shot alone
    flat_k = torch.from_numpy(synthetic.rec_flat_index([rec[k]], nz, nx)).cuda()
    rk = ar.propagate(vp, rho, b, w, s[k:k + 1].contiguous(), flat_k, 1, dt, h)
    counts = [len(r) for r in rec]
    off = sum(counts[:k])
    same = torch.equal(rk, r1[:, off:off + counts[k]])
    print(f"rho full size: linearity {e_lin:.2e}, shot {k} alone == in batch: {same}")
    assert e_lin < 1e-7 and same
