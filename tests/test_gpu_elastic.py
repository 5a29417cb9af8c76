"""Parity of the HIP elastic path against the FP64 oracle and the FP32 torch
restatement of the reference path (same seeded inputs).

Tolerances (relative L2) as for the acoustic path: 1e-5 against the FP32
restatement where its own FP32 error against FP64 is small, otherwise no
worse than 2x the restatement's own error against the FP64 oracle.
"""
import numpy as np
import pytest
import torch

from oracle import elastic as oe
from oracle import ref_torch
from tests.cases import elastic_case, elastic_dev, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _oracle_ok(e_hip, e_ref):
    return e_hip <= 2.0 * e_ref + 1e-6


def test_single_step_matches_reference():
    from seistorch_amd.equations2d import elastic as el
    c = elastic_case()
    rng = np.random.default_rng(0)
    f = [rng.standard_normal((2, c.nz, c.nx)).astype(np.float32) for _ in range(5)]
    t = elastic_dev(c)
    out = el._time_step(t["vp"], t["vs"], t["rho"], *[torch.from_numpy(x).cuda() for x in f],
                        torch.tensor(c.dt), torch.tensor(c.h), t["d"])
    ref = ref_torch.elastic_step(torch.from_numpy(c.vp), torch.from_numpy(c.vs),
                                 torch.from_numpy(c.rho), *[torch.from_numpy(x) for x in f],
                                 torch.tensor(c.dt), torch.tensor(c.h), torch.from_numpy(c.d))
    k = oe.coefficients(c.vp.astype(np.float64), c.vs.astype(np.float64),
                        c.rho.astype(np.float64), c.d.astype(np.float64), c.dt, c.h)
    o64, _ = oe.step(k, *[x.astype(np.float64) for x in f])
    errs = [rel_l2(a.cpu().numpy(), b.numpy()) for a, b in zip(out, ref)]
    errs64 = [rel_l2(a.cpu().numpy(), b) for a, b in zip(out, o64)]
    print("elastic step rel-L2 vs ref_torch:", ["%.1e" % e for e in errs],
          "vs FP64:", ["%.1e" % e for e in errs64])
    assert max(errs) < 1e-6 and max(errs64) < 1e-5


def test_single_step_vjp_matches_autograd():
    from seistorch_amd.equations2d import elastic as el
    c = elastic_case()
    rng = np.random.default_rng(1)
    f = [rng.standard_normal((2, c.nz, c.nx)) for _ in range(5)]
    gy = [rng.standard_normal((2, c.nz, c.nx)) for _ in range(5)]
    m64 = [torch.from_numpy(x.astype(np.float64)).requires_grad_(True) for x in (c.vp, c.vs, c.rho)]
    f64 = [torch.from_numpy(x).requires_grad_(True) for x in f]
    out = ref_torch.elastic_step(*m64, *f64, torch.tensor(c.dt, dtype=torch.float64),
                                 torch.tensor(c.h, dtype=torch.float64),
                                 torch.from_numpy(c.d.astype(np.float64)))
    torch.autograd.backward(out, [torch.from_numpy(g) for g in gy])
    t = elastic_dev(c)
    m = [t[k].clone().requires_grad_(True) for k in ("vp", "vs", "rho")]
    fh = [torch.from_numpy(x.astype(np.float32)).cuda().requires_grad_(True) for x in f]
    outh = el._time_step(*m, *fh, c.dt, c.h, t["d"])
    torch.autograd.backward(outh, [torch.from_numpy(g.astype(np.float32)).cuda() for g in gy])
    errs = [rel_l2(a.grad.cpu(), b.grad) for a, b in zip(fh + m, f64 + m64)]
    print("elastic vjp rel-L2 (vx vz txx tzz txz vp vs rho):", ["%.1e" % e for e in errs])
    assert max(errs) < 1e-5


def test_reverse_step_inverts_forward_on_interior():
    """_time_step_backward (reference reverse step with rings) undoes a step.
    The state is a physical wavefield (80 steps from an explosive source), so
    stresses and velocities carry their impedance ratio."""
    from seistorch_amd.equations2d import elastic as el
    from seistorch_amd.equations2d.utils import save_boundaries
    c = elastic_case(nshots=2)
    t = elastic_dev(c)
    s0 = [torch.zeros((2, c.nz, c.nx), device="cuda") for _ in range(5)]
    bidx = torch.arange(2, device="cuda")
    sz, sx = t["src"][:, 0].long(), t["src"][:, 1].long()
    for n in range(80):
        s0 = list(el._time_step(t["vp"], t["vs"], t["rho"], *s0, c.dt, c.h, t["d"]))
        for k in (2, 3):
            s0[k][bidx, sz, sx] += t["w"][n]
    s1 = el._time_step(t["vp"], t["vs"], t["rho"], *s0, c.dt, c.h, t["d"])
    rings = [save_boundaries(x) for x in s0]
    no_src = lambda y, w, dt=1.0: y  # noqa: E731
    back = el._time_step_backward(t["vp"], t["vs"], t["rho"], *[x.detach() for x in s1], c.dt,
                                  c.h, t["d"], rings, [["txx"], no_src, 0.0])
    I = (slice(None), slice(50, -50), slice(50, -50))
    errs = [rel_l2(a.detach().cpu().numpy()[I], b.cpu().numpy()[I]) for a, b in zip(back, s0)]
    print("reverse-step rel-L2 per field:", ["%.1e" % e for e in errs])
    assert max(errs) < 1e-5


def _records_fused(c, t, src_type=("txx", "tzz"), rec_type=("vx", "vz"), vp=None, vs=None,
                   rho=None, S=None, K=None):
    from seistorch_amd.equations2d import elastic as el
    return el.propagate(t["vp"] if vp is None else vp, t["vs"] if vs is None else vs,
                        t["rho"] if rho is None else rho, t["d"], t["w"], t["src"], t["rec"],
                        c.B, c.dt, c.h, src_type, rec_type, shots_per_group=S, snapshot_every=K)


@pytest.mark.parametrize("S", [None, 1, 8])
def test_fused_forward_records(S):
    c = elastic_case()
    t = elastic_dev(c)
    rec = _records_fused(c, t, S=S).cpu().numpy()
    ro = np.concatenate(oe.forward(c.vp, c.vs, c.rho, c.d, c.dt, c.h, c.w, c.src, c.rec), 1)
    rr = torch.cat(ref_torch.elastic_forward(c.vp, c.vs, c.rho, c.d, c.dt, c.h, c.w, c.src,
                                             c.rec), 1).numpy()
    e_or, e_ref, e_ro = rel_l2(rec, ro), rel_l2(rec, rr), rel_l2(rr, ro)
    print(f"elastic records S={S}: HIP vs oracle {e_or:.2e}, HIP vs ref_torch {e_ref:.2e}, "
          f"ref_torch vs oracle {e_ro:.2e}")
    assert e_ref < TOL and _oracle_ok(e_or, e_ro)


@pytest.mark.parametrize("K", [None, 0, 9])
def test_fused_adjoint_interior(K):
    c = elastic_case(nt=250)
    t = elastic_dev(c)
    ro = oe.forward(c.vp, c.vs, c.rho, c.d, c.dt, c.h, c.w, c.src, c.rec)
    syn = np.concatenate(ro, 1)
    rng = np.random.default_rng(3)
    G = (rng.standard_normal(syn.shape) * np.abs(syn).max()).astype(np.float32)
    m = [t[k].clone().requires_grad_(True) for k in ("vp", "vs", "rho")]
    rec = _records_fused(c, t, vp=m[0], vs=m[1], rho=m[2], K=K)
    rec.backward(torch.from_numpy(G).cuda())
    gh = [x.grad.cpu().numpy() for x in m]
    obs_o, o = [], 0
    for x in ro:
        n = x.shape[1]
        obs_o.append(x - 0.5 * G[:, o:o + n].astype(np.float64))
        o += n
    _, go, _ = oe.gradient(c.vp, c.vs, c.rho, c.d, c.dt, c.h, c.w, c.src, c.rec, obs_o)
    _, gr, _ = ref_torch.elastic_gradient_ad(c.vp, c.vs, c.rho, c.d, c.dt, c.h, c.w, c.src,
                                             c.rec, None, rec_cotangent=G)
    I = (slice(50, -50), slice(50, -50))
    for name, a, b, r in zip(("vp", "vs", "rho"), gh, go, gr):
        r = r.numpy()
        e_or, e_ref, e_ro = rel_l2(a[I], b[I]), rel_l2(a[I], r[I]), rel_l2(r[I], b[I])
        print(f"elastic adjoint K={K} d{name}: HIP vs oracle {e_or:.2e}, HIP vs ref_torch AD "
              f"{e_ref:.2e}, ref_torch AD vs oracle {e_ro:.2e}")
        # both FP32 paths carry ~9e-6 against FP64 at nt=300 (measured: HIP vs
        # the restatement 8.2e-6..9.6e-6 over snapshot spacings 0, 9 and the
        # default): bounded at twice that, and at twice the restatement's own
        # error against FP64
        assert _oracle_ok(e_or, e_ro) and e_ref < 2 * TOL
        mask = np.ones_like(a, bool); mask[I] = False
        assert np.all(a[mask] == 0.0)


def test_gradient_finite_difference_fp32():
    """Central differences of J = sum(G * records) at random interior cells."""
    c = elastic_case(nz_i=30, nx_i=40, nshots=2, nt=200)
    t = elastic_dev(c)
    rng = np.random.default_rng(4)
    R = sum(len(r) for r in c.rec)
    G = torch.from_numpy(rng.standard_normal((c.nt, R, 2)).astype(np.float32)).cuda()
    m = [t[k].clone().requires_grad_(True) for k in ("vp", "vs", "rho")]
    rec = _records_fused(c, t, vp=m[0], vs=m[1], rho=m[2])
    rec.backward(G)
    for which, step in ((0, 10.0), (1, 10.0), (2, 10.0)):
        g = m[which].grad.cpu().numpy()
        for _ in range(2):
            i = int(rng.integers(60, c.nz - 55)); j = int(rng.integers(55, c.nx - 55))
            vals = []
            for sgn in (1.0, -1.0):
                mm = [t[k].clone() for k in ("vp", "vs", "rho")]
                mm[which][i, j] += sgn * step
                r = _records_fused(c, t, vp=mm[0], vs=mm[1], rho=mm[2])
                vals.append(float((r.double() * G.double()).sum()))
            fd = (vals[0] - vals[1]) / (2 * step)
            print(f"elastic FD param {which} cell ({i},{j}): fd {fd:.6e} adjoint {g[i, j]:.6e}")
            assert abs(fd - g[i, j]) <= 2e-2 * max(abs(fd), 1e-2 * np.abs(g).max())


@pytest.mark.parametrize("nx_i", [61, 62, 60])
def test_lane_vector_widths(nx_i):
    """Widths whose padded nx takes 1-, 2- and 4-cell lane vectors: records
    bit-identical to the FP32 restatement, gradients as in the adjoint test."""
    c = elastic_case(nx_i=nx_i, nt=150)
    t = elastic_dev(c)
    m = [t[k].clone().requires_grad_(True) for k in ("vp", "vs", "rho")]
    rec = _records_fused(c, t, vp=m[0], vs=m[1], rho=m[2])
    rr = torch.cat(ref_torch.elastic_forward(c.vp, c.vs, c.rho, c.d, c.dt, c.h, c.w, c.src,
                                             c.rec), 1).numpy()
    assert rel_l2(rec.detach().cpu().numpy(), rr) == 0.0
    rng = np.random.default_rng(nx_i)
    G = (rng.standard_normal(rr.shape) * np.abs(rr).max()).astype(np.float32)
    rec.backward(torch.from_numpy(G).cuda())
    _, gr, _ = ref_torch.elastic_gradient_ad(c.vp, c.vs, c.rho, c.d, c.dt, c.h, c.w, c.src,
                                             c.rec, None, rec_cotangent=G)
    I = (slice(50, -50), slice(50, -50))
    for name, a, r in zip(("vp", "vs", "rho"), m, gr):
        e = rel_l2(a.grad.cpu().numpy()[I], r.numpy()[I])
        print(f"elastic nx={c.nx} d{name}: HIP vs ref_torch AD {e:.2e}")
        assert e < 2 * TOL


@pytest.mark.parametrize("nt,S,nx_i", [(120, None, 60), (121, 1, 62), (1, 2, 61), (2, 8, 60)])
def test_fused_step_matches_two_stages(nt, S, nx_i):
    """One pass per forward step and per reverse step (neighbouring rows
    recomputed, lane shuffles, ping-pong buffers) gives bit-identical records
    and bit-identical gradients."""
    c = elastic_case(nx_i=nx_i, nt=nt, delay=min(40, nt))
    t = elastic_dev(c)
    rng = np.random.default_rng(nt)
    out = {}
    for fused in (False, True):
        m = [t[k].clone().requires_grad_(True) for k in ("vp", "vs", "rho")]
        from seistorch_amd.equations2d import elastic as el
        rec = el.propagate(m[0], m[1], m[2], t["d"], t["w"], t["src"], t["rec"], c.B, c.dt, c.h,
                           shots_per_group=S, fused=fused)
        G = torch.from_numpy(rng.standard_normal(tuple(rec.shape)).astype(np.float32)).cuda() \
            if fused is False else out["G"]
        out["G"] = G
        (rec * G).sum().backward()
        out[fused] = (rec.detach().cpu().numpy(), [x.grad.cpu().numpy() for x in m])
    assert np.array_equal(out[False][0], out[True][0])
    for name, a, b in zip(("vp", "vs", "rho"), out[True][1], out[False][1]):
        assert np.array_equal(a, b), (name, float(np.abs(a - b).max()), float(np.abs(b).max()))
    print(f"elastic fused step nt={nt} S={S} nx={c.nx}: records equal, gradients equal")
