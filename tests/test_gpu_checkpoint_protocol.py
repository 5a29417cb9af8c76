"""The per-step drop-in functions driven the way the reference's
boundary-saving checkpoint drives them (seistorch/checkpoint_new.py:114-217,
called from cell.py:72-73 inside the rnn.py:178-212 loop):

* forward: each step runs ``_time_step`` without grad, the rings of both
  outputs are kept (save_boundaries, equations2d/utils.py:50-69), then the
  source is added and the receivers are read;
* backward, one step at a time from the last: ``_time_step_backward`` with
  (h1, h2) = (u^n, u^{n+1}), the ring of u^{n-1} and the source function
  rebuilds u^{n-1}; the step (u^n, u^{n-1}) -> u^{n+1} is recomputed with
  grad and the cotangents are pushed through it (its VJP kernel).

The loop follows the exact time pairing (the reference's own loop shifts the
wavelet by three samples and skips a step, a defect of that path that is not
reproduced).  The gradient on the interior is compared with the fused
adjoint (``propagate``) and with FP64 autograd; the rebuild is ring-only, as
the reference's, so the FP32 bar is the reference path's own boundary-saving
error (oracle/ref_torch.gradient_bs, FP32 vs FP64).
"""
import importlib

import numpy as np  # noqa: F401
import pytest
import torch

from oracle import ref_torch
from seistorch_amd.equations2d.utils import save_boundaries
from tests.cases import acoustic_case, rel_l2, to_dev
from tests.test_gpu_compile import PointSource, _smask

pytestmark = pytest.mark.gpu

I = (slice(50, -50), slice(50, -50))


def _protocol_gradient(mod, vp, damp, c, t, G, kw):
    B, nz, nx = c.B, c.nz, c.nx
    src = PointSource(_smask(c.src, B, nz, nx))
    w = t["w"]
    rec = t["rec"]
    u = torch.zeros((B, nz, nx), device="cuda")
    up = torch.zeros_like(u)
    rings = []
    records = []
    with torch.no_grad():
        for n in range(c.nt):
            y, h1 = mod._time_step(vp, u, up, c.dt, c.h, damp, **kw)
            rings.append((save_boundaries(y), save_boundaries(h1)))
            up = h1
            u = src(y, w[n])
            records.append(u.reshape(-1)[rec])
    grad = torch.zeros_like(vp)
    u_next, u_cur = u, up           # u^{n+1}, u^n for n = nt-1
    a_next = torch.zeros_like(u)    # dJ/du^{n+1}
    a_cur = torch.zeros_like(u)     # contributions to dJ/du^n so far
    for n in range(c.nt - 1, -1, -1):
        a_next = a_next.reshape(-1).index_add(0, rec, G[n]).reshape(B, nz, nx)
        with torch.no_grad():
            if n >= 2:
                # rings of the outputs of step n-2, whose first output is u^{n-1}
                prev, _ = mod._time_step_backward(vp, u_cur, u_next, c.dt, c.h, damp,
                                                  rings[n - 2], [["p"], src, w[n]])
            else:
                prev = torch.zeros_like(u)  # u^0 = u^-1 = 0
        vpg = vp.detach().clone().requires_grad_(True)
        uc = u_cur.detach().clone().requires_grad_(True)
        upv = prev.detach().clone().requires_grad_(True)
        y, _ = mod._time_step(vpg, uc, upv, c.dt, c.h, damp, **kw)
        y.backward(a_next)
        grad += vpg.grad
        a_next, a_cur = a_cur + uc.grad, upv.grad
        u_next, u_cur = u_cur, prev
    return grad, torch.stack(records)


@pytest.mark.parametrize("equation", ["acoustic", "acoustic_habc"])
def test_per_step_backward_in_checkpoint_protocol(equation):
    mod = importlib.import_module(f"seistorch_amd.equations2d.{equation}")
    c = acoustic_case(nz_i=50, nx_i=70, nshots=2, nt=300, seed=4)
    t = to_dev(c)
    habc = equation == "acoustic_habc"
    if habc:
        from seistorch_amd.habc import generate_habc_coefficients_2d
        damp = generate_habc_coefficients_2d((c.nz, c.nx), 50).cuda()
        kw = {"habcs": [torch.ones(1)] * 4}
    else:
        damp = t["d"]
        kw = {"habcs": None}
    R = t["rec"].numel()
    g = torch.Generator(device="cuda")
    g.manual_seed(9)
    G = torch.randn((c.nt, R), generator=g, device="cuda")
    gp, rp = _protocol_gradient(mod, t["vp"], damp, c, t, G, kw)
    # the fused loop and its adjoint
    vp = t["vp"].clone().requires_grad_(True)
    rf = mod.propagate(vp, damp, t["w"], t["src"], t["rec"], c.B, c.dt, c.h)
    rf.backward(G)
    gf = vp.grad.cpu().numpy()
    gp = gp.cpu().numpy()
    Gn = G.cpu().numpy()
    dn = damp.cpu().numpy()
    if habc:
        g64, _ = ref_torch.habc_gradient_ad(c.vp, dn, c.dt, c.h, c.w, c.src, c.rec, Gn,
                                            dtype=torch.float64)
    else:
        _, g64, _ = ref_torch.gradient_ad(c.vp, dn, c.dt, c.h, c.w, c.src, c.rec, None,
                                          dtype=torch.float64, rec_cotangent=Gn)
    _, gbs, _ = ref_torch.gradient_bs(c.vp, dn, c.dt, c.h, c.w, c.src, c.rec, None,
                                      rec_cotangent=Gn, habc=habc)
    g64, gbs = g64.numpy(), gbs.numpy()
    e_pf = rel_l2(gp[I], gf[I])
    e_p64 = rel_l2(gp[I], g64[I])
    e_f64 = rel_l2(gf[I], g64[I])
    e_bs64 = rel_l2(gbs[I], g64[I])
    print(f"{equation} checkpoint protocol: records equal fused {torch.equal(rp, rf.detach())}; "
          f"gradient protocol vs fused {e_pf:.2e}, protocol vs FP64 {e_p64:.2e}, fused vs FP64 "
          f"{e_f64:.2e}, FP32 BS restatement vs FP64 {e_bs64:.2e}", flush=True)
    assert torch.equal(rp, rf.detach())
    assert e_p64 <= 2 * e_bs64 + 1e-6
    # outside the rebuilt interior the protocol's gradient is not a gradient
    # (the reference compares [50:-50] only, AD_vs_BS.py:7-11): the reverse
    # step of a damped cell amplifies, so only the interior is checked
    assert np.isfinite(gp[I]).all()
