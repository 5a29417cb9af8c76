"""Parity of the HIP HABC path against the FP32 torch strip restatement of the
reference (oracle/ref_torch.habc_*), the FP64 per-cell oracle (oracle/habc)
and FP64 autograd of the restatement for the transposes."""
import numpy as np
import pytest
import torch

from oracle import habc as oh
from oracle import ref_torch
from tests.cases import acoustic_case, rel_l2, to_dev

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _habc_case(nshots=3, nt=300, multiple=False, **kw):
    from seistorch_amd.habc import generate_habc_coefficients_2d
    c = acoustic_case(nshots=nshots, nt=nt, **kw)
    c.b = generate_habc_coefficients_2d((c.nz, c.nx), 50, multiple=multiple).numpy()
    c.multiple = multiple
    return c


def _masks(c):
    from seistorch_amd.habc import bound_mask
    return bound_mask(c.nz, c.nx, 50, "cuda", c.B, return_idx=True, multiple=c.multiple)


@pytest.mark.parametrize("multiple", [False, True])
def test_step_matches_reference(multiple):
    from seistorch_amd.equations2d import acoustic_habc as ah
    c = _habc_case(multiple=multiple)
    rng = np.random.default_rng(0)
    h1 = rng.standard_normal((2, c.nz, c.nx)).astype(np.float32)
    h2 = rng.standard_normal((2, c.nz, c.nx)).astype(np.float32)
    y, _ = ah._time_step(torch.from_numpy(c.vp).cuda(), torch.from_numpy(h1).cuda(),
                         torch.from_numpy(h2).cuda(), torch.tensor(c.dt), torch.tensor(c.h),
                         torch.from_numpy(c.b).cuda(), habcs=_masks(c))
    yr, _ = ref_torch.habc_step(torch.from_numpy(c.vp), torch.from_numpy(h1), torch.from_numpy(h2),
                                torch.tensor(c.dt), torch.tensor(c.h), torch.from_numpy(c.b),
                                multiple=multiple)
    yo = oh.step(c.vp.astype(np.float64), h1.astype(np.float64), h2.astype(np.float64),
                 c.b.astype(np.float64), c.dt, c.h, multiple)
    e_ref, e_or = rel_l2(y.cpu().numpy(), yr.numpy()), rel_l2(y.cpu().numpy(), yo)
    print(f"habc step multiple={multiple}: vs ref_torch {e_ref:.2e}, vs FP64 {e_or:.2e}")
    assert e_ref < 1e-6 and e_or < 1e-6


@pytest.mark.parametrize("multiple", [False, True])
def test_step_vjp_matches_autograd(multiple):
    from seistorch_amd.equations2d import acoustic_habc as ah
    c = _habc_case(multiple=multiple)
    rng = np.random.default_rng(1)
    h1, h2, gy = (rng.standard_normal((2, c.nz, c.nx)) for _ in range(3))
    vp64 = torch.from_numpy(c.vp.astype(np.float64)).requires_grad_(True)
    a1 = torch.from_numpy(h1).requires_grad_(True)
    a2 = torch.from_numpy(h2).requires_grad_(True)
    y, _ = ref_torch.habc_step(vp64, a1, a2, torch.tensor(c.dt, dtype=torch.float64),
                               torch.tensor(c.h, dtype=torch.float64),
                               torch.from_numpy(c.b.astype(np.float64)), multiple=multiple)
    y.backward(torch.from_numpy(gy))
    vp = torch.from_numpy(c.vp).cuda().requires_grad_(True)
    b1 = torch.from_numpy(h1.astype(np.float32)).cuda().requires_grad_(True)
    b2 = torch.from_numpy(h2.astype(np.float32)).cuda().requires_grad_(True)
    yh, _ = ah._time_step(vp, b1, b2, c.dt, c.h, torch.from_numpy(c.b).cuda(), habcs=_masks(c))
    yh.backward(torch.from_numpy(gy.astype(np.float32)).cuda())
    errs = [rel_l2(b1.grad.cpu(), a1.grad), rel_l2(b2.grad.cpu(), a2.grad),
            rel_l2(vp.grad.cpu(), vp64.grad)]
    print(f"habc vjp multiple={multiple} (h1, h2, vp):", ["%.2e" % e for e in errs])
    assert max(errs) < 1e-5


def _fused(c, t, vp=None, K=None, S=None):
    from seistorch_amd.equations2d import acoustic_habc as ah
    return ah.propagate(t["vp"] if vp is None else vp, torch.from_numpy(c.b).cuda(), t["w"],
                        t["src"], t["rec"], c.B, c.dt, c.h, c.multiple, shots_per_group=S,
                        snapshot_every=K)


@pytest.mark.parametrize("S,nx_i", [(None, 90), (1, 90), (None, 540), (2, 263)])
def test_fused_records(S, nx_i):
    """Records of the fused loop; nx_i 540 and 263 have deep interiors run by
    the vectorised launch (4- and 1-cell lanes): still bit-identical."""
    c = _habc_case(nshots=4, nx_i=nx_i, nt=300 if nx_i == 90 else 200)
    t = to_dev(c)
    rec = _fused(c, t, S=S).cpu().numpy()
    rr = torch.cat(ref_torch.habc_forward(c.vp, c.b, c.dt, c.h, c.w, c.src, c.rec), 1)[..., 0].numpy()
    e_ref = rel_l2(rec, rr)
    if nx_i == 90:
        ro = np.concatenate(oh.forward(c.vp, c.b, c.dt, c.h, c.w, c.src, c.rec), 1)[..., 0]
        e_or, e_ro = rel_l2(rec, ro), rel_l2(rr, ro)
        print(f"habc records S={S}: vs ref_torch {e_ref:.2e}, vs FP64 {e_or:.2e}, ref vs FP64 {e_ro:.2e}")
        assert e_or <= 2 * e_ro + 1e-6
    print(f"habc records S={S} grid {c.nz}x{c.nx}: vs ref_torch {e_ref:.2e}")
    assert e_ref == 0.0


@pytest.mark.parametrize("K,nx_i,nz_i", [(None, 90, 60), (0, 90, 60), (None, 540, 70),
                                         (0, 263, 66)])
def test_fused_adjoint_interior(K, nx_i, nz_i):
    """nx_i 540 (padded nx 640, 4-cell lanes) and 263 (363, 1-cell lanes) give
    tiles deep inside the strips, which take the vectorised adjoint path."""
    c = _habc_case(nshots=3, nt=200, nx_i=nx_i, nz_i=nz_i)
    t = to_dev(c)
    rr = ref_torch.habc_forward(c.vp, c.b, c.dt, c.h, c.w, c.src, c.rec)
    syn = torch.cat(rr, 1)[..., 0].numpy()
    rng = np.random.default_rng(2)
    G = (rng.standard_normal(syn.shape) * np.abs(syn).max()).astype(np.float32)
    vp = t["vp"].clone().requires_grad_(True)
    rec = _fused(c, t, vp=vp, K=K)
    rec.backward(torch.from_numpy(G).cuda())
    g = vp.grad.cpu().numpy()
    g32, _ = ref_torch.habc_gradient_ad(c.vp, c.b, c.dt, c.h, c.w, c.src, c.rec, G)
    g64, _ = ref_torch.habc_gradient_ad(c.vp, c.b, c.dt, c.h, c.w, c.src, c.rec, G,
                                        dtype=torch.float64)
    I = (slice(50, -50), slice(50, -50))
    e_ref, e_64, e_r64 = (rel_l2(g[I], g32.numpy()[I]), rel_l2(g[I], g64.numpy()[I]),
                          rel_l2(g32.numpy()[I], g64.numpy()[I]))
    print(f"habc adjoint K={K} grid {c.nz}x{c.nx}: vs ref_torch AD {e_ref:.2e}, vs FP64 AD {e_64:.2e}, "
          f"ref vs FP64 {e_r64:.2e}")
    assert e_ref < TOL and e_64 <= 2 * e_r64 + 1e-6
    mask = np.ones_like(g, bool); mask[I] = False
    assert np.all(g[mask] == 0.0)
