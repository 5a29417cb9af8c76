"""Free surface (the reference's geom.multiple) on the fused loops.

With `multiple` the model is padded at the bottom and sides only
(geom.py:213-240), the damping has no top strip (pml.py:49-53, habc.py:4-81)
and the reference's _time_step_backward_multiple leaves the top ring row alone
(equations2d/utils.py:71-78), so the reverse-time rebuild -- and with it the
boundary-saving gradient -- covers rows [0, nz-50) of the interior columns
(acoustic.py:114-137, acoustic_habc.py:247-267, elastic.py:124-207).

Bar: the gradient on that region within 1e-5 of autograd through the FP32
restatement of the reference loop (the same bar as without a free surface),
no worse than 2x the restatement's own error against FP64 autograd, exactly
zero outside it, and not zero in the top rows.  Without snapshots (ring-only
rebuild, the reference's own footprint) the rebuild drifts in FP32 over the
whole reversed loop, as the reference's boundary-saving path does: that case
is held to twice the FP32 boundary-saving restatement's error against FP64.
"""
import numpy as np
import pytest
import torch

from oracle import ref_torch
from seistorch_amd import synthetic
from seistorch_amd.pml import generate_pml_coefficients_2d
from seistorch_amd.utils import pad_model, ricker_wave
from tests.cases import Case, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5
P = 50


def fs_case(nz_i=60, nx_i=90, nshots=3, nt=200, dt=1e-3, h=10.0, fm=15.0, seed=11,
            src_depth=2, elastic=False):
    """A free-surface problem: top rows are model rows, sources at src_depth,
    receivers on row 0 (grid indices without the top shift, setup.py:386-483)."""
    c = Case()
    v = synthetic.layered_model(nz_i, nx_i, 1800.0 if elastic else 1500.0, 3000.0, seed=seed)
    c.vp = pad_model(v, multiple=True).astype(np.float32)
    c.nz, c.nx = c.vp.shape
    c.d = generate_pml_coefficients_2d((c.nz, c.nx), P, multiple=True).numpy()
    c.w = ricker_wave(fm, dt, nt, delay=60).numpy()
    src, rec = synthetic.surface_acquisition(nx_i, nshots, src_depth=src_depth)
    c.src = src - np.array([P, 0])
    c.rec = [r - np.array([P, 0]) for r in rec]
    c.dt, c.h, c.nt, c.B = dt, h, nt, nshots
    if elastic:
        vs = v / 1.73
        vs[:6] = 0.0
        c.vs = pad_model(vs, multiple=True).astype(np.float32)
        c.rho = pad_model(310.0 * v ** 0.25, multiple=True).astype(np.float32)
    return c


def dev(c):
    t = {k: torch.from_numpy(np.ascontiguousarray(getattr(c, k))).cuda()
         for k in ("vp", "d", "w")}
    t["src"] = torch.from_numpy(c.src.astype(np.int32)).cuda()
    t["rec"] = torch.from_numpy(synthetic.rec_flat_index(c.rec, c.nz, c.nx)).cuda()
    return t


REGION = (slice(0, -P), slice(P, -P))


def _check_region(name, g, g32, g64, g_bs=None, tol=TOL):
    """g_bs: the FP32 boundary-saving restatement of the reference path
    (ring-only reverse-time rebuild, oracle/ref_torch.gradient_bs), given when
    g comes from a ring-only rebuild too (no snapshots).  Both rebuilds drift
    from the forward in FP32 over the whole reversed loop, so that case is
    held to no more than twice the reference path's own BS error against FP64
    instead of the 1e-5 bar against autograd."""
    e_ref = rel_l2(g[REGION], g32[REGION])
    e_64 = rel_l2(g[REGION], g64[REGION])
    e_r64 = rel_l2(g32[REGION], g64[REGION])
    top = np.abs(g[:P, P:-P]).max() / np.abs(g[REGION]).max()
    msg = (f"{name}: region rel-L2 vs FP32 AD {e_ref:.2e}, vs FP64 AD {e_64:.2e}, FP32 AD vs "
           f"FP64 AD {e_r64:.2e}")
    if g_bs is not None:
        e_bs64 = rel_l2(g_bs[REGION], g64[REGION])
        msg += f", FP32 BS restatement vs FP64 AD {e_bs64:.2e}"
    print(f"{msg}; top rows peak / region peak {top:.2e}", flush=True)
    if g_bs is None:
        assert e_ref < tol and e_64 <= 2 * e_r64 + 1e-6
    else:
        assert e_64 <= 2 * e_bs64 + 1e-6
    assert top > 1e-3                      # the top rows carry gradient
    mask = np.ones_like(g, bool)
    mask[REGION] = False
    assert np.all(g[mask] == 0.0)


def _cotangent(syn, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(syn.shape) * np.abs(syn).max()).astype(np.float32)


@pytest.mark.parametrize("K", [None, 0])
def test_acoustic_pml_free_surface_gradient(K):
    from seistorch_amd.equations2d import acoustic as ac
    c = fs_case()
    t = dev(c)
    rr = ref_torch.forward(c.vp, c.d, c.dt, c.h, c.w, c.src, c.rec)
    syn = torch.cat(rr, 1)[..., 0].numpy()
    G = _cotangent(syn, 1)
    vp = t["vp"].clone().requires_grad_(True)
    rec = ac.propagate(vp, t["d"], t["w"], t["src"], t["rec"], c.B, c.dt, c.h,
                       snapshot_every=K, multiple=True)
    assert torch.equal(rec.detach().cpu(), torch.from_numpy(syn))
    rec.backward(torch.from_numpy(G).cuda())
    g = vp.grad.cpu().numpy()
    _, g32, _ = ref_torch.gradient_ad(c.vp, c.d, c.dt, c.h, c.w, c.src, c.rec, None,
                                      rec_cotangent=G)
    _, g64, _ = ref_torch.gradient_ad(c.vp, c.d, c.dt, c.h, c.w, c.src, c.rec, None,
                                      dtype=torch.float64, rec_cotangent=G)
    g_bs = None
    if K == 0:
        g_bs = ref_torch.gradient_bs(c.vp, c.d, c.dt, c.h, c.w, c.src, c.rec, None,
                                     multiple=True, rec_cotangent=G)[1].numpy()
    _check_region(f"acoustic PML free surface K={K}", g, g32.numpy(), g64.numpy(), g_bs)


def test_acoustic_free_surface_finite_difference():
    """FP32 central differences of J = sum(G * records) at cells of the top
    rows, which only the free-surface region covers."""
    from seistorch_amd.equations2d import acoustic as ac
    c = fs_case(nz_i=40, nx_i=60, nshots=2, nt=300)
    t = dev(c)
    rng = np.random.default_rng(7)
    G = torch.from_numpy(rng.standard_normal((c.nt, 2 * 60)).astype(np.float32)).cuda()
    vp = t["vp"].clone().requires_grad_(True)
    rec = ac.propagate(vp, t["d"], t["w"], t["src"], t["rec"], c.B, c.dt, c.h, multiple=True)
    rec.backward(G)
    g = vp.grad.cpu().numpy()
    for i, j in ((1, 70), (10, 80), (30, 95)):
        vals = []
        for sgn in (1.0, -1.0):
            vpp = t["vp"].clone()
            vpp[i, j] += sgn * 10.0
            r = ac.propagate(vpp, t["d"], t["w"], t["src"], t["rec"], c.B, c.dt, c.h,
                             multiple=True)
            vals.append(float((r.double() * G.double()).sum()))
        fd = (vals[0] - vals[1]) / 20.0
        print(f"free-surface FD cell ({i},{j}): fd {fd:.6e} adjoint {g[i, j]:.6e}", flush=True)
        assert abs(fd - g[i, j]) <= 1e-2 * max(abs(fd), 1e-2 * np.abs(g).max())


@pytest.mark.parametrize("K", [None, 0])
def test_habc_free_surface_gradient(K):
    from seistorch_amd.equations2d import acoustic_habc as ah
    from seistorch_amd.habc import generate_habc_coefficients_2d
    c = fs_case()
    b = generate_habc_coefficients_2d((c.nz, c.nx), P, multiple=True).numpy()
    t = dev(c)
    rr = ref_torch.habc_forward(c.vp, b, c.dt, c.h, c.w, c.src, c.rec, multiple=True)
    syn = torch.cat(rr, 1)[..., 0].numpy()
    G = _cotangent(syn, 2)
    vp = t["vp"].clone().requires_grad_(True)
    rec = ah.propagate(vp, torch.from_numpy(b).cuda(), t["w"], t["src"], t["rec"], c.B, c.dt,
                       c.h, multiple=True, snapshot_every=K)
    assert torch.equal(rec.detach().cpu(), torch.from_numpy(syn))
    rec.backward(torch.from_numpy(G).cuda())
    g = vp.grad.cpu().numpy()
    g32, _ = ref_torch.habc_gradient_ad(c.vp, b, c.dt, c.h, c.w, c.src, c.rec, G, multiple=True)
    g64, _ = ref_torch.habc_gradient_ad(c.vp, b, c.dt, c.h, c.w, c.src, c.rec, G,
                                        dtype=torch.float64, multiple=True)
    g_bs = None
    if K == 0:
        g_bs = ref_torch.gradient_bs(c.vp, b, c.dt, c.h, c.w, c.src, c.rec, None, multiple=True,
                                     rec_cotangent=G, habc=True)[1].numpy()
    _check_region(f"HABC free surface K={K}", g, g32.numpy(), g64.numpy(), g_bs)


def test_elastic_free_surface_gradient():
    from seistorch_amd.equations2d import elastic as el
    c = fs_case(nz_i=40, nx_i=60, nt=250, dt=5e-4, h=5.0, fm=20.0, elastic=True)
    t = dev(c)
    m = [torch.from_numpy(getattr(c, k)).cuda().requires_grad_(True) for k in ("vp", "vs", "rho")]
    rr = ref_torch.elastic_forward(c.vp, c.vs, c.rho, c.d, c.dt, c.h, c.w, c.src, c.rec)
    syn = torch.cat(rr, 1).numpy()
    G = _cotangent(syn, 3)
    rec = el.propagate(m[0], m[1], m[2], t["d"], t["w"], t["src"], t["rec"], c.B, c.dt, c.h,
                       multiple=True, snapshot_every=25)
    assert torch.equal(rec.detach().cpu(), torch.from_numpy(syn))
    rec.backward(torch.from_numpy(G).cuda())
    _, g32, _ = ref_torch.elastic_gradient_ad(c.vp, c.vs, c.rho, c.d, c.dt, c.h, c.w, c.src,
                                              c.rec, None, rec_cotangent=G)
    _, g64, _ = ref_torch.elastic_gradient_ad(c.vp, c.vs, c.rho, c.d, c.dt, c.h, c.w, c.src,
                                              c.rec, None, dtype=torch.float64, rec_cotangent=G)
    # as for the elastic gradients without a free surface: the FP32 sums of
    # the coefficient gradients over steps and shots (and their chain to vp,
    # vs, rho) round differently from autograd's; this 250-step case lands
    # at 8.5e-6 against FP64 (autograd 4.5e-6), the two 1.03e-5 apart for vp,
    # with snapshots every 25 or 50 steps alike: held to 2e-5 against FP32
    # autograd and twice autograd's own error against FP64
    for name, x, a, b in zip(("vp", "vs", "rho"), m, g32, g64):
        _check_region(f"elastic free surface d{name}", x.grad.cpu().numpy(), a.numpy(),
                      b.numpy(), tol=2 * TOL)


def test_free_surface_needs_undamped_top():
    """A gradient with multiple=True over a damping that has a top strip would
    rebuild damped rows: refused before any launch."""
    from seistorch_amd.equations2d import acoustic as ac
    c = fs_case(nt=20)
    t = dev(c)
    d_top = torch.from_numpy(generate_pml_coefficients_2d((c.nz, c.nx), P).numpy()).cuda()
    vp = t["vp"].clone().requires_grad_(True)
    with pytest.raises(RuntimeError, match="free surface"):
        ac.propagate(vp, d_top, t["w"], t["src"], t["rec"], c.B, c.dt, c.h, multiple=True)


def test_non_finite_records_raise():
    """rnn.py:215: records with NaN/inf raise ValueError (dt far beyond CFL)."""
    from seistorch_amd.equations2d import acoustic as ac
    c = fs_case(nt=400, dt=2e-2)
    t = dev(c)
    with pytest.raises(ValueError, match="non-finite"):
        ac.propagate(t["vp"], t["d"], t["w"], t["src"], t["rec"], c.B, c.dt, c.h)
