"""The drop-in equation functions under the reference's forced torch.compile.

seistorch compiles both plug-in functions unconditionally
(seistorch/compile.py:3,10-13: ``torch.compile(func)`` with default
arguments) before handing them to its cell (seistorch/model.py:83-87, which
also sets an ``ACOUSTIC2nd`` attribute on the compiled forward).  Here the
same wrapping is applied to the HIP-backed functions; the compiled calls
must give what the eager calls give -- the per-step operators run the same
kernels either way -- for the forward step, the reverse step with rings and
a source function, and the gradient of a short autograd loop through the
compiled step (the reference's boundary_saving: false path, cell.py:74-75).
"""
import numpy as np
import pytest
import torch

from seistorch_amd.equations2d.utils import save_boundaries
from tests.cases import acoustic_case, elastic_case, rel_l2, to_dev, elastic_dev

pytestmark = pytest.mark.gpu


class PointSource(torch.nn.Module):
    """The part of seistorch/source.py:47-57 the backward functions call:
    ``src_func(Y, X, dt)`` returns a copy of Y with smask * X added."""

    def __init__(self, smask):
        super().__init__()
        self.register_buffer("smask", smask)

    def forward(self, Y, X, dt=1.0):
        Y_new = Y.clone()
        Y_new += self.smask * X
        return Y_new


def _compile(fn):
    # seistorch/compile.py:10-13 with force_compile = True
    return torch.compile(fn)


def _smask(src, B, nz, nx):
    m = torch.zeros((B, nz, nx), device="cuda")
    for b, (z, x) in enumerate(src):
        m[b, z, x] = 1.0
    return m


def _fields(shape, seed, n):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return [torch.randn(shape, generator=g, device="cuda") for _ in range(n)]


@pytest.mark.parametrize("equation", ["acoustic", "acoustic_habc"])
def test_compiled_second_order_steps_equal_eager(equation):
    import importlib
    mod = importlib.import_module(f"seistorch_amd.equations2d.{equation}")
    c = acoustic_case(nz_i=30, nx_i=40, nshots=2, nt=10)
    t = to_dev(c)
    if equation == "acoustic_habc":
        from seistorch_amd.habc import generate_habc_coefficients_2d
        damp = generate_habc_coefficients_2d((c.nz, c.nx), 50).cuda()
        kw = {"habcs": [torch.ones(1)] * 4}
    else:
        damp = t["d"]
        kw = {"habcs": None}
    fwd, bwd = _compile(mod._time_step), _compile(mod._time_step_backward)
    fwd.ACOUSTIC2nd = True  # model.py:87
    h1, h2 = _fields((c.B, c.nz, c.nx), 1, 2)
    ye, pe = mod._time_step(t["vp"], h1, h2, c.dt, c.h, damp, **kw)
    yc, pc = fwd(t["vp"], h1, h2, c.dt, c.h, damp, **kw)
    assert torch.equal(ye, yc) and torch.equal(pe, pc)
    # reverse step: (h1, h2) = (u^n, u^{n+1}), ring of the rebuilt field, source re-added
    src = PointSource(_smask(c.src, c.B, c.nz, c.nx))
    rings = (save_boundaries(h2), save_boundaries(h1))
    w = torch.tensor(0.37, device="cuda")
    be, _ = mod._time_step_backward(t["vp"], h1, h2, c.dt, c.h, damp, rings, [["p"], src, w])
    bc, _ = bwd(t["vp"], h1, h2, c.dt, c.h, damp, rings, [["p"], src, w])
    assert torch.equal(be, bc)
    # fullgraph: the forward step traces without a graph break
    ff = torch.compile(mod._time_step, fullgraph=True)
    yf, _ = ff(t["vp"], h1, h2, c.dt, c.h, damp, **kw)
    assert torch.equal(ye, yf)


@pytest.mark.parametrize("equation", ["acoustic", "acoustic_habc"])
def test_autograd_loop_through_compiled_step(equation):
    """Records and dJ/dvp of a 60-step loop (step -> source -> receivers,
    rnn.py:178-212) through the compiled step equal the eager loop's."""
    import importlib
    mod = importlib.import_module(f"seistorch_amd.equations2d.{equation}")
    c = acoustic_case(nz_i=30, nx_i=40, nshots=2, nt=60)
    t = to_dev(c)
    if equation == "acoustic_habc":
        from seistorch_amd.habc import generate_habc_coefficients_2d
        damp = generate_habc_coefficients_2d((c.nz, c.nx), 50).cuda()
        kw = {"habcs": [torch.ones(1)] * 4}
    else:
        damp = t["d"]
        kw = {"habcs": None}
    smask = _smask(c.src, c.B, c.nz, c.nx)
    w = t["w"]
    rec = t["rec"]
    G = _fields((c.nt, rec.numel()), 5, 1)[0]

    def run(step):
        vp = t["vp"].clone().requires_grad_(True)
        u = torch.zeros((c.B, c.nz, c.nx), device="cuda")
        up = torch.zeros_like(u)
        out = []
        for n in range(c.nt):
            y, up = step(vp, u, up, c.dt, c.h, damp, **kw)
            u = y.clone()
            u += smask * w[n]
            out.append(u.reshape(-1)[rec])
        r = torch.stack(out)
        (r * G).sum().backward()
        return r.detach(), vp.grad

    re_, ge = run(mod._time_step)
    rc, gc = run(_compile(mod._time_step))
    e_r, e_g = rel_l2(rc.cpu().numpy(), re_.cpu().numpy()), rel_l2(gc.cpu().numpy(),
                                                                  ge.cpu().numpy())
    print(f"{equation}: compiled vs eager loop: records {e_r:.2e}, gradient {e_g:.2e}",
          flush=True)
    assert e_r <= 1e-7 and e_g <= 1e-6


def test_compiled_elastic_steps_equal_eager():
    from seistorch_amd.equations2d import elastic as el
    c = elastic_case(nz_i=30, nx_i=40, nshots=2, nt=10)
    t = elastic_dev(c)
    f = _fields((c.B, c.nz, c.nx), 2, 5)
    args = (t["vp"], t["vs"], t["rho"], *f, c.dt, c.h, t["d"])
    oe = el._time_step(*args)
    oc = _compile(el._time_step)(*args)
    assert all(torch.equal(a, b) for a, b in zip(oe, oc))
    src = PointSource(_smask(c.src, c.B, c.nz, c.nx))
    rings = [save_boundaries(x) for x in f]
    w = torch.tensor(0.25, device="cuda")
    be = el._time_step_backward(*args, rings, [["txx", "tzz"], src, w])
    bc = _compile(el._time_step_backward)(*args, rings, [["txx", "tzz"], src, w])
    assert all(torch.equal(a, b) for a, b in zip(be, bc))
    # gradient through the compiled step equals the eager one
    gs = []
    for fn in (el._time_step, _compile(el._time_step)):
        m = [x.clone().requires_grad_(True) for x in (t["vp"], t["vs"], t["rho"])]
        out = fn(*m, *f, c.dt, c.h, t["d"])
        sum((o * g).sum() for o, g in zip(out, _fields((c.B, c.nz, c.nx), 3, 5))).backward()
        gs.append([x.grad for x in m])
    for a, b in zip(*gs):
        assert rel_l2(b.cpu().numpy(), a.cpu().numpy()) <= 1e-6
