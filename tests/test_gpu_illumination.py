"""Source illumination accumulated by the fused acoustic forward
(rnn.py:127-128, 204-205: zeroed per forward, then per time level the sum
over shots of the squared post-source field is added), the divisor of
process.py:114-117."""
import numpy as np
import pytest
import torch

from oracle import acoustic as oa
from tests.cases import acoustic_case, rel_l2, to_dev

pytestmark = pytest.mark.gpu


def _per_step_illumination(c, t):
    """The reference's loop on the drop-in per-step HIP function, with the
    illumination summed in FP32 torch ops in the reference's order."""
    from seistorch_amd.equations2d import acoustic as ac
    B = c.B
    smask = torch.zeros((B, c.nz, c.nx), device="cuda")
    smask[torch.arange(B), t["src"][:, 0].long(), t["src"][:, 1].long()] = 1.0
    h1 = torch.zeros_like(smask)
    h2 = torch.zeros_like(smask)
    P = torch.zeros((c.nz, c.nx), device="cuda")
    for n in range(c.nt):
        h1, h2 = ac._time_step(t["vp"], h1, h2, c.dt, c.h, t["d"])
        h1 = h1 + smask * t["w"][n]
        s = torch.zeros_like(P)
        for b in range(B):
            s = s + h1[b] * h1[b]
        P = P + s
    return P


@pytest.mark.parametrize("nt,two_step,S", [(200, True, None), (201, True, None),
                                           (201, False, None), (200, True, 1), (7, True, 4)])
def test_illumination_matches_the_reference_loop(nt, two_step, S):
    from seistorch_amd.equations2d import acoustic as ac
    c = acoustic_case(nshots=3, nt=nt)
    t = to_dev(c)
    illum = torch.full((c.nz, c.nx), 7.0, device="cuda")  # overwritten, not accumulated
    rec = ac.propagate(t["vp"], t["d"], t["w"], t["src"], t["rec"], c.B, c.dt, c.h, S,
                       two_step=two_step, illum=illum)
    rec0 = ac.propagate(t["vp"], t["d"], t["w"], t["src"], t["rec"], c.B, c.dt, c.h, S,
                        two_step=two_step)
    assert torch.equal(rec, rec0)
    ref32 = _per_step_illumination(c, t)
    e32 = rel_l2(illum.cpu().numpy(), ref32.cpu().numpy())
    bitwise = torch.equal(illum, ref32)
    _, fields = oa.forward(c.vp, c.d, c.dt, c.h, c.w, c.src, c.rec, keep_fields=True)
    ref64 = sum((u ** 2).sum(0) for u in fields[1:])
    e64 = rel_l2(illum.cpu().numpy(), ref64)
    print(f"illumination nt={nt} two_step={two_step} S={S}: vs FP32 reference loop {e32:.2e} "
          f"(bitwise {bitwise}), vs FP64 oracle {e64:.2e}")
    if two_step and nt >= 2:
        # the two-step launches sum per shot group and the groups at the end:
        # the same nt*B positive terms in another order, so within the
        # worst-case relative error of such a sum in FP32
        assert e32 <= nt * c.B * 2.0 ** -24
    else:
        assert bitwise
    assert e64 < 1e-5


def test_illumination_preconditioned_gradient():
    """propagate(illum=) -> gradient -> PostProcess.precondition: the
    preconditioned gradient is the plain one divided by the illumination."""
    from seistorch_amd import process as pp
    from seistorch_amd.equations2d import acoustic as ac
    c = acoustic_case(nshots=2, nt=200)
    t = to_dev(c)
    illum = torch.empty((c.nz, c.nx), device="cuda")
    vp = torch.nn.Parameter(t["vp"].clone())
    rec = ac.propagate(vp, t["d"], t["w"], t["src"], t["rec"], c.B, c.dt, c.h, illum=illum)
    (rec ** 2).sum().backward()
    g0 = vp.grad.clone()
    post = pp.PostProcess(pp.Params([vp], precondition=illum), {}, type("C", (), {})())
    post.precondition()
    # cells the wavefield never reached hold 0 / 0, as in the reference
    assert torch.allclose(vp.grad, g0 / illum, rtol=0, atol=0, equal_nan=True)


def test_habc_illumination_matches_the_reference_loop():
    """The HABC forward's illumination against the reference loop on the
    per-step HABC function, FP32 sums in the reference's order."""
    from seistorch_amd.equations2d import acoustic_habc as ah
    from seistorch_amd.habc import bound_mask, generate_habc_coefficients_2d
    c = acoustic_case(nshots=3, nt=150)
    t = to_dev(c)
    b = generate_habc_coefficients_2d((c.nz, c.nx), 50).cuda()
    illum = torch.empty((c.nz, c.nx), device="cuda")
    rec = ah.propagate(t["vp"], b, t["w"], t["src"], t["rec"], c.B, c.dt, c.h, illum=illum)
    rec0 = ah.propagate(t["vp"], b, t["w"], t["src"], t["rec"], c.B, c.dt, c.h)
    assert torch.equal(rec, rec0)
    masks = bound_mask(c.nz, c.nx, 50, "cuda", c.B, return_idx=True)
    smask = torch.zeros((c.B, c.nz, c.nx), device="cuda")
    smask[torch.arange(c.B), t["src"][:, 0].long(), t["src"][:, 1].long()] = 1.0
    h1 = torch.zeros_like(smask)
    h2 = torch.zeros_like(smask)
    P = torch.zeros((c.nz, c.nx), device="cuda")
    for n in range(c.nt):
        h1, h2 = ah._time_step(t["vp"], h1, h2, c.dt, c.h, b, habcs=masks)
        h1 = h1 + smask * t["w"][n]
        s = torch.zeros_like(P)
        for k in range(c.B):
            s = s + h1[k] * h1[k]
        P = P + s
    e = rel_l2(illum.cpu().numpy(), P.cpu().numpy())
    print(f"HABC illumination vs the per-step loop: {e:.2e} (bitwise {torch.equal(illum, P)})")
    assert e < 1e-6


@pytest.mark.parametrize("fused", [True, False])
def test_elastic_illumination_matches_the_reference_loop(fused):
    """The elastic forward's illumination of the last source type's field
    (rnn.py:197-205) against the reference loop on the per-step function."""
    from seistorch_amd.equations2d import elastic as el
    from tests.cases import elastic_case, elastic_dev
    c = elastic_case(nshots=2, nt=120)
    t = elastic_dev(c)
    src_type = ("txx", "tzz")
    illum = torch.empty((c.nz, c.nx), device="cuda")
    rec = el.propagate(t["vp"], t["vs"], t["rho"], t["d"], t["w"], t["src"], t["rec"], c.B,
                       c.dt, c.h, src_type=src_type, fused=fused, illum=illum)
    rec0 = el.propagate(t["vp"], t["vs"], t["rho"], t["d"], t["w"], t["src"], t["rec"], c.B,
                        c.dt, c.h, src_type=src_type, fused=fused)
    assert torch.equal(rec, rec0)
    smask = torch.zeros((c.B, c.nz, c.nx), device="cuda")
    smask[torch.arange(c.B), t["src"][:, 0].long(), t["src"][:, 1].long()] = 1.0
    f = [torch.zeros_like(smask) for _ in range(5)]
    P = torch.zeros((c.nz, c.nx), device="cuda")
    k = el.FIELDS.index(src_type[-1])
    for n in range(c.nt):
        f = list(el._time_step(t["vp"], t["vs"], t["rho"], *f, c.dt, c.h, t["d"]))
        for st in src_type:
            j = el.FIELDS.index(st)
            f[j] = f[j] + smask * t["w"][n]
        s = torch.zeros_like(P)
        for b in range(c.B):
            s = s + f[k][b] * f[k][b]
        P = P + s
    e = rel_l2(illum.cpu().numpy(), P.cpu().numpy())
    print(f"elastic illumination fused={fused} vs the per-step loop: {e:.2e} "
          f"(bitwise {torch.equal(illum, P)})")
    assert e < 1e-6
