"""The torch.library operators are registered with fake implementations and
autograd formulas (no GPU needed: fake tensors never reach the library)."""
import pytest
import torch
from torch._subclasses.fake_tensor import FakeTensorMode

from seistorch_amd import library  # noqa: F401  (registers the operators)

OPS = ("acoustic_step", "acoustic_step_vjp", "habc_step", "habc_step_vjp", "elastic_stage",
       "elastic_stage_vjp", "rho_habc_step")


@pytest.mark.parametrize("name", OPS)
def test_operator_registered(name):
    op = getattr(torch.ops.seistorch_amd, name)
    assert op.default._schema.name == f"seistorch_amd::{name}"


def test_fake_implementations_give_output_shapes():
    B, nz, nx = 3, 12, 17
    with FakeTensorMode() as mode:
        vp = mode.from_tensor(torch.ones(nz, nx))
        u = mode.from_tensor(torch.ones(B, nz, nx))
        y = torch.ops.seistorch_amd.acoustic_step(vp, u, u, vp, 1e-3, 10.0)
        assert tuple(y.shape) == (B, nz, nx)
        g = torch.ops.seistorch_amd.acoustic_step_vjp(vp, vp, u, u, 1e-3, 10.0)
        assert [tuple(t.shape) for t in g] == [(nz, nx), (B, nz, nx), (B, nz, nx)]
        y = torch.ops.seistorch_amd.habc_step(vp, u, u, vp, 1e-3, 10.0, False)
        assert tuple(y.shape) == (B, nz, nx)
        f = torch.ops.seistorch_amd.elastic_stage(vp, vp, vp, vp, u, u, u, u, u, 0, 1.0, 0, nz,
                                                  0, nx, 5e-4, 5.0)
        assert len(f) == 5 and all(tuple(t.shape) == (B, nz, nx) for t in f)
        g = torch.ops.seistorch_amd.elastic_stage_vjp(vp, vp, vp, vp, u, u, u, u, u, u, u, u, u,
                                                      u, 1, -1.0, 1, nz - 1, 1, nx - 1, 5e-4,
                                                      5.0)
        assert [tuple(t.shape) for t in g] == [(nz, nx)] * 3 + [(B, nz, nx)] * 5
        y = torch.ops.seistorch_amd.rho_habc_step(vp, vp, u, u, None, 1e-3, 10.0, False, False)
        assert tuple(y.shape) == (B, nz, nx)


def test_real_call_on_host_tensors_raises():
    """The operators run HIP kernels only: host tensors are refused."""
    vp = torch.ones(12, 17)
    u = torch.ones(2, 12, 17)
    with pytest.raises(RuntimeError, match="HIP device"):
        torch.ops.seistorch_amd.acoustic_step(vp, u, u, vp, 1e-3, 10.0)
