"""Uneven multi-model batch: every instance must reproduce the single-instance
fused loop (records bit-identical; gradients equal up to the order of the
receiver-cell atomic adds, which is not fixed even between two single runs),
with direct launches and inside one hipGraph, replayed with eager work in
between."""
import numpy as np
import pytest
import torch

from tests.cases import acoustic_case, rel_l2, to_dev

pytestmark = pytest.mark.gpu

# (nz_i, nx_i, shots, nt, dt, h, fm): different grids, shot counts, time axes
# (two share nt, one stops after 2 steps, one after 1, one odd), dt and h
SPECS = [(60, 92, 3, 300, 1e-3, 10.0, 15.0),
         (40, 60, 2, 120, 8e-4, 8.0, 20.0),
         (70, 80, 5, 300, 1e-3, 10.0, 12.0),
         (30, 40, 1, 2, 1e-3, 10.0, 15.0),
         (50, 52, 4, 1, 1e-3, 10.0, 15.0),
         (36, 100, 6, 180, 6e-4, 6.0, 25.0),
         (44, 64, 3, 151, 1e-3, 10.0, 15.0)]
GRAD_TOL = 1e-6


def _instances(seed=0):
    from seistorch_amd.equations2d.acoustic_batch import BatchInstance
    out, weights = [], []
    rng = np.random.default_rng(seed)
    for q, (nz_i, nx_i, B, nt, dt, h, fm) in enumerate(SPECS):
        c = acoustic_case(nz_i=nz_i, nx_i=nx_i, nshots=B, nt=nt, dt=dt, h=h, fm=fm, seed=q + 1,
                          delay=min(60, nt))
        t = to_dev(c)
        out.append(BatchInstance(t["vp"], t["d"], t["w"], t["src"], t["rec"], B, dt, h))
        weights.append(torch.from_numpy(
            rng.standard_normal((nt, t["rec"].shape[0])).astype(np.float32)).cuda())
    return out, weights


def _batch(insts, weights, **kw):
    from seistorch_amd.equations2d.acoustic_batch import propagate_batch
    for it in insts:
        it.vp = it.vp.detach().clone().requires_grad_(True)
    recs = propagate_batch(insts, **kw)
    loss = sum((r * w).sum() for r, w in zip(recs, weights))
    loss.backward()
    torch.cuda.synchronize()
    return [r.detach().clone() for r in recs], [it.vp.grad.clone() for it in insts]


def _single(insts, weights, S, rz, vx, K):
    from seistorch_amd.equations2d import acoustic as ac
    recs, grads = [], []
    for it, w in zip(insts, weights):
        v = it.vp.detach().clone().requires_grad_(True)
        r = ac.propagate(v, it.d, it.wavelet, it.src_zx, it.rec_flat, it.B, it.dt, it.h,
                         shots_per_group=S, snapshot_every=K, rows_per_thread=rz, vec_x=vx)
        (r * w).sum().backward()
        recs.append(r.detach().clone())
        grads.append(v.grad.clone())
    torch.cuda.synchronize()
    return recs, grads


@pytest.mark.parametrize("S,rz,vx,K", [(4, 2, 4, 0), (2, 1, 1, 50), (8, 4, 2, 0), (1, 2, 4, 60)])
def test_batch_matches_single_instances(S, rz, vx, K):
    insts, weights = _instances()
    rb, gb = _batch(insts, weights, shots_per_group=S, rows_per_thread=rz, vec_x=vx,
                    snapshot_every=K, use_graph=False)
    rs, gs = _single(insts, weights, S, rz, vx, K)
    for q, (a, b, ga, gc) in enumerate(zip(rb, rs, gb, gs)):
        assert a.shape == b.shape, q
        assert torch.equal(a, b), f"instance {q}: records differ"
        e = rel_l2(ga.cpu().numpy(), gc.cpu().numpy())
        print(f"instance {q}: grad rel-L2 batch vs single {e:.2e}")
        assert e <= GRAD_TOL, (q, e)
        assert float(gc.abs().max()) > 0 or SPECS[q][3] <= 2


def test_batch_against_the_oracle():
    """Every instance of one batched launch against the FP32 restatement of
    the reference loop (oracle/ref_torch: records bit for bit, the gradient of
    sum(w * records) by autograd through the loop) and the FP64 oracle
    (oracle/acoustic records, ref_torch autograd in FP64), on the rebuilt
    interior: no worse than twice the restatement's own FP64 error."""
    from oracle import acoustic as oa
    from oracle import ref_torch
    insts, weights = _instances(seed=5)
    rb, gb = _batch(insts, weights, shots_per_group=2, rows_per_thread=2, vec_x=4,
                    snapshot_every=None, use_graph=False)
    I = (slice(50, -50), slice(50, -50))
    for q, (nz_i, nx_i, B, nt, dt, h, fm) in enumerate(SPECS):
        c = acoustic_case(nz_i=nz_i, nx_i=nx_i, nshots=B, nt=nt, dt=dt, h=h, fm=fm, seed=q + 1,
                          delay=min(60, nt))
        rr = torch.cat(ref_torch.forward(c.vp, c.d, c.dt, c.h, c.w, c.src, c.rec), 1)[..., 0]
        ro = np.concatenate(oa.forward(c.vp, c.d, c.dt, c.h, c.w, c.src, c.rec), 1)[..., 0]
        r = rb[q].cpu().numpy()
        e_ref, e_or, e_ro = rel_l2(r, rr.numpy()), rel_l2(r, ro), rel_l2(rr.numpy(), ro)
        w = weights[q].cpu().numpy()
        _, g32, _ = ref_torch.gradient_ad(c.vp, c.d, c.dt, c.h, c.w, c.src, c.rec, None,
                                          rec_cotangent=w)
        _, g64, _ = ref_torch.gradient_ad(c.vp, c.d, c.dt, c.h, c.w, c.src, c.rec, None,
                                          dtype=torch.float64, rec_cotangent=w)
        g = gb[q].cpu().numpy()[I]
        g32, g64 = g32.numpy()[I], g64.numpy()[I]
        eg, eg64, eg_r = rel_l2(g, g32), rel_l2(g, g64), rel_l2(g32, g64)
        print(f"instance {q} ({c.nz}x{c.nx}, {B} shots, nt={nt}): records vs restatement "
              f"{e_ref:.2e}, vs FP64 {e_or:.2e} (restatement {e_ro:.2e}); gradient vs FP32 AD "
              f"{eg:.2e}, vs FP64 {eg64:.2e} (FP32 AD {eg_r:.2e})", flush=True)
        assert e_ref == 0.0 and e_or <= 2 * e_ro + 1e-6
        if nt > 2:  # nt <= 2: no adjoint step reaches the interior gradient
            assert eg < 1e-5 and eg64 <= 2 * eg_r + 1e-6


def test_graph_replays_match_direct_launches():
    insts, weights = _instances(seed=3)
    shape = dict(shots_per_group=4, rows_per_thread=2, vec_x=4, snapshot_every=50)
    r0, g0 = _batch(insts, weights, use_graph=False, **shape)
    junk = torch.zeros(1 << 20, device="cuda")
    for rep in range(3):
        r1, g1 = _batch(insts, weights, use_graph=True, **shape)
        junk.add_(1.0)  # eager work between replays
        torch.cuda.synchronize()
        for q in range(len(insts)):
            assert torch.equal(r0[q], r1[q]), (rep, q)
            assert rel_l2(g1[q].cpu().numpy(), g0[q].cpu().numpy()) <= GRAD_TOL, (rep, q)


@pytest.mark.parametrize("S,rz,vx,K,graph", [(4, 2, 4, 50, False), (2, 1, 1, 0, True),
                                             (8, 4, 2, 60, False)])
def test_two_step_forward_matches_one_step(S, rz, vx, K, graph):
    """The batched forward at two time steps per launch (four rotating field
    arenas; instances whose last step falls in the first half of a pair stop
    there) against one step per launch: records bit for bit, and the
    gradients, which read the forward's final fields, rings and snapshots."""
    insts, weights = _instances()
    out = {two: _batch(insts, weights, shots_per_group=S, rows_per_thread=rz, vec_x=vx,
                       snapshot_every=K, use_graph=graph, two_step=two) for two in (True, False)}
    for q in range(len(insts)):
        assert torch.equal(out[True][0][q], out[False][0][q]), f"instance {q}: records differ"
        e = rel_l2(out[True][1][q].cpu().numpy(), out[False][1][q].cpu().numpy())
        assert e <= GRAD_TOL, (q, e)


def test_forward_only_batch_has_no_grad_state():
    from seistorch_amd.equations2d.acoustic_batch import propagate_batch
    insts, _ = _instances()
    with torch.no_grad():
        rb = propagate_batch(insts, shots_per_group=4, rows_per_thread=2, vec_x=4)
    rs, _ = _single(insts, [torch.zeros(1, device="cuda")] * len(insts), 4, 2, 4, 0)
    for a, b in zip(rb, rs):
        assert torch.equal(a, b)


def test_batch_rejects_bad_inputs():
    from seistorch_amd.equations2d.acoustic_batch import propagate_batch
    insts, _ = _instances()
    with pytest.raises(RuntimeError, match="vec_x"):
        propagate_batch(insts[:2] + [_odd_width()], vec_x=4, shots_per_group=2,
                        rows_per_thread=1)
    bad = insts[1]
    bad.src_zx = bad.src_zx.clone()
    bad.src_zx[0, 1] = bad.vp.shape[1]
    with pytest.raises(RuntimeError, match="instance 1"):
        propagate_batch(insts[:2])


def _odd_width():
    from seistorch_amd.equations2d.acoustic_batch import BatchInstance
    c = acoustic_case(nz_i=30, nx_i=43, nshots=2, nt=20)
    t = to_dev(c)
    return BatchInstance(t["vp"], t["d"], t["w"], t["src"], t["rec"], 2, c.dt, c.h)
