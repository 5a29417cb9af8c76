"""Extra point sources in one field (source encoding, codingfwi.py:184-334:
each encoded source adds its own wavelet sample at its cell,
source.py:54-55): records and the adjoint gradient of the fused acoustic loop
against autograd through the reference-order loop with the same sources in
FP64 and FP32 (the step of oracle/ref_torch.time_step, pinned in
tests/test_oracle_acoustic.py), and superposition against one run per
source."""
import numpy as np
import pytest
import torch

from oracle import ref_torch
from tests.cases import acoustic_case, rel_l2, to_dev

pytestmark = pytest.mark.gpu


def _encoded(c, n_src, seed):
    """Source cells (z, x) and wavelets (n_src, nt): random polarity and time
    shift of the case's wavelet, as codingfwi's roll() encodes them."""
    rng = np.random.default_rng(seed)
    xs = rng.choice(np.arange(55, c.nx - 55), n_src, replace=False)
    zs = rng.integers(52, 60, n_src)
    wav = np.zeros((n_src, c.nt), np.float32)
    for k in range(n_src):
        tau = int(rng.integers(0, int(0.2 * c.nt)))
        sign = -1.0 if rng.integers(1, 3) % 2 else 1.0
        wav[k, tau:] = sign * c.w[: c.nt - tau]
    return np.stack([zs, xs], 1).astype(np.int64), wav


def _ad_loop(c, cells, wav, G, dtype=torch.float64):
    """Records and dJ/dc of J = sum(G * records) by autograd through the
    reference-order loop (step, then every source's sample), one field."""
    dt = torch.tensor(c.dt, dtype=dtype)
    h = torch.tensor(c.h, dtype=dtype)
    cp = torch.from_numpy(c.vp).to(dtype).requires_grad_(True)
    d = torch.from_numpy(c.d).to(dtype)
    rec = torch.as_tensor(np.asarray(c.rec[0]), dtype=torch.int64)
    u = torch.zeros((1, c.nz, c.nx), dtype=dtype)
    up = torch.zeros_like(u)
    w = torch.from_numpy(wav).to(dtype)
    out = []
    for n in range(c.nt):
        y, up = ref_torch.time_step(cp, u, up, dt, h, d)
        add = torch.zeros_like(y)
        for k, (z, x) in enumerate(cells):
            add[0, z, x] = add[0, z, x] + w[k, n]
        u = y + add
        out.append(u[0, rec[:, 0], rec[:, 1]])
    r = torch.stack(out)
    (r * torch.from_numpy(G).to(dtype)).sum().backward()
    return r.detach().numpy(), cp.grad.numpy()


@pytest.mark.parametrize("K", [None, 0])
def test_encoded_sources_against_fp64(K):
    from seistorch_amd.equations2d import acoustic as ac
    c = acoustic_case(nz_i=50, nx_i=90, nshots=1, nt=300, seed=6)
    cells, wav = _encoded(c, 5, 1)
    t = to_dev(c)
    src = torch.from_numpy(cells[:1].astype(np.int32)).cuda()
    extra = (torch.from_numpy(np.concatenate([np.zeros((4, 1), np.int64), cells[1:]], 1)
                              .astype(np.int32)).cuda(),
             torch.from_numpy(wav[1:]).cuda())
    G = np.random.default_rng(2).standard_normal((c.nt, len(c.rec[0]))).astype(np.float32)
    vp = t["vp"].clone().requires_grad_(True)
    r = ac.propagate(vp, t["d"], torch.from_numpy(wav[0]).cuda(), src, t["rec"], 1, c.dt, c.h,
                     extra_sources=extra, snapshot_every=K)
    r.backward(torch.from_numpy(G).cuda())
    r64, g64 = _ad_loop(c, cells, wav, G)
    r32, g32 = _ad_loop(c, cells, wav, G, torch.float32)
    I = (slice(50, -50), slice(50, -50))
    er, er_ref = rel_l2(r.detach().cpu().numpy(), r64), rel_l2(r32, r64)
    eg, eg_ref = rel_l2(vp.grad.cpu().numpy()[I], g64[I]), rel_l2(g32[I], g64[I])
    e32 = rel_l2(vp.grad.cpu().numpy()[I], g32[I])
    print(f"encoded sources (5 in one field, snapshot_every={K}): records vs FP64 {er:.2e} "
          f"(FP32 loop {er_ref:.2e}); gradient vs FP64 {eg:.2e} (FP32 autograd {eg_ref:.2e}), "
          f"vs FP32 autograd {e32:.2e}", flush=True)
    assert er <= 2 * er_ref + 1e-6
    if K == 0:  # ring-only rebuild: FP32 drift over the whole reversed loop
        assert eg <= 10 * eg_ref + 1e-6
    else:
        assert eg <= 2 * eg_ref + 1e-6 and e32 < 1e-5


def test_encoded_sources_superpose():
    """The one-field run equals the sum of one run per source (FP32, up to
    rounding); a source table outside the grid is refused."""
    from seistorch_amd.equations2d import acoustic as ac
    c = acoustic_case(nz_i=50, nx_i=90, nshots=1, nt=250, seed=7)
    cells, wav = _encoded(c, 4, 3)
    t = to_dev(c)
    src0 = torch.from_numpy(cells[:1].astype(np.int32)).cuda()
    es = np.concatenate([np.zeros((3, 1), np.int64), cells[1:]], 1).astype(np.int32)
    with torch.no_grad():
        r_all = ac.propagate(t["vp"], t["d"], torch.from_numpy(wav[0]).cuda(), src0, t["rec"], 1,
                             c.dt, c.h, extra_sources=(torch.from_numpy(es).cuda(),
                                                       torch.from_numpy(wav[1:]).cuda()))
        parts = [ac.propagate(t["vp"], t["d"], torch.from_numpy(wav[k]).cuda(),
                              torch.from_numpy(cells[k:k + 1].astype(np.int32)).cuda(), t["rec"],
                              1, c.dt, c.h) for k in range(4)]
    e = rel_l2(r_all.cpu().numpy(), sum(p.cpu().numpy() for p in parts))
    print(f"encoded sources superposition: {e:.2e}", flush=True)
    # FP32 superposition holds to rounding and gradual underflow (see the
    # linearity check of the full-size FWI configuration)
    assert e < 1e-5
    bad = es.copy()
    bad[0, 1] = c.nz
    with pytest.raises(RuntimeError, match="extra source"):
        ac.propagate(t["vp"], t["d"], torch.from_numpy(wav[0]).cuda(), src0, t["rec"], 1, c.dt,
                     c.h, extra_sources=(torch.from_numpy(bad).cuda(),
                                         torch.from_numpy(wav[1:]).cuda()))


def test_encoded_fwi_loop_lowers_the_misfit():
    """run_encoded (codingfwi.py:184-334): per epoch all shots encoded into one
    field, one forward and one adjoint.  The codes change every epoch, so the
    epochs are compared through the plain misfit of all shots, which must
    fall.  The model RMSE is printed only: a few unsmoothed updates of this
    small problem fit the shallow acquisition footprint first (55.9 -> 61.1
    m/s over 8 epochs while the misfit falls 2.7x), as plain FWI does."""
    from seistorch_amd import fwi
    from tests.test_gpu_learned import _small_problem
    p = _small_problem()
    m0 = p.misfit(p.vp_init)
    r0 = p.rmse(p.vp_init)
    vp, hist = fwi.run_encoded(p, ["all"], epochs=8, lr=10.0, log=print)
    m1, r1 = p.misfit(vp), p.rmse(vp)
    print(f"encoded FWI: plain misfit {m0:.4e} -> {m1:.4e}, rmse {r0:.2f} -> {r1:.2f}",
          flush=True)
    assert len(hist.misfit) == 8 and all(np.isfinite(hist.misfit))
    assert m1 < 0.8 * m0 and np.isfinite(r1)
