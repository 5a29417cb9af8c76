"""The per-step HIP kernels as ``torch.library`` operators.

seistorch wraps every equation's ``_time_step`` and ``_time_step_backward``
in ``torch.compile`` unconditionally (seistorch/compile.py:3,10-13, applied in
seistorch/model.py:83-85).  A ctypes call inside a traced function cannot be
traced, so the drop-in modules call these operators instead: each one has a
fake (shape-only) implementation for tracing and its hand-written VJP kernel
registered as its autograd formula, so the compiled functions keep calling
the HIP kernels and differentiate through them.

    torch.ops.seistorch_amd.acoustic_step(vp, h1, h2, d, dt, h)            -> y
    torch.ops.seistorch_amd.habc_step(vp, h1, h2, b, dt, h, multiple)      -> y
    torch.ops.seistorch_amd.elastic_stage(vp, vs, rho, d, vx, vz, txx, tzz, txz,
                                          stage, sign, z0, z1, x0, x1, dt, h)
                                                                           -> 5 fields
    torch.ops.seistorch_amd.rho_habc_step(vp, rho, p1, p2, b, dt, h, multiple, blend)
                                                                           -> y

Each takes FP32 tensors on a HIP device and launches on the current stream;
a failing call raises RuntimeError.
"""
from typing import Optional

import torch

from . import _lib

NS = "seistorch_amd"


def _check_pair(name, vp, h1, h2, m):
    _lib.require_device(vp, h1, h2, m)
    if h1.shape != h2.shape or h1.dim() != 3 or vp.shape != h1.shape[1:] or m.shape != vp.shape:
        raise RuntimeError(f"{name}: shape mismatch vp{tuple(vp.shape)} h1{tuple(h1.shape)} "
                           f"h2{tuple(h2.shape)} {tuple(m.shape)}")


# ------------------------------------------------------------- acoustic (PML)

@torch.library.custom_op(f"{NS}::acoustic_step", mutates_args=())
def acoustic_step(vp: torch.Tensor, h1: torch.Tensor, h2: torch.Tensor, d: torch.Tensor,
                  dt: float, h: float) -> torch.Tensor:
    """y = A*(2/dt^2*h1 - B*h2 + vp^2*L h1) (seistorch/equations2d/acoustic.py:81-84)."""
    h1, h2 = h1.contiguous(), h2.contiguous()
    _check_pair("acoustic_step", vp, h1, h2, d)
    B, nz, nx = h1.shape
    y = torch.empty_like(h1)
    _lib.call("seis_acoustic_step", _lib.ptr(vp), _lib.ptr(d), _lib.ptr(h1), _lib.ptr(h2),
              _lib.ptr(y), B, nz, nx, dt, h, _lib.stream_of(h1))
    return y


@acoustic_step.register_fake
def _(vp, h1, h2, d, dt, h):
    return torch.empty_like(h1, memory_format=torch.contiguous_format)


@torch.library.custom_op(f"{NS}::acoustic_step_vjp", mutates_args=())
def acoustic_step_vjp(vp: torch.Tensor, d: torch.Tensor, h1: torch.Tensor, gy: torch.Tensor,
                      dt: float, h: float) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """(dJ/dvp summed over shots, dJ/dh1, dJ/dh2) for dJ/dy = gy."""
    h1, gy = h1.contiguous(), gy.contiguous()
    B, nz, nx = h1.shape
    gh1 = torch.empty_like(h1)
    gh2 = torch.empty_like(h1)
    gvp = torch.empty_like(vp)
    _lib.call("seis_acoustic_step_vjp", _lib.ptr(vp), _lib.ptr(d), _lib.ptr(h1), _lib.ptr(gy),
              None, _lib.ptr(gh1), _lib.ptr(gh2), _lib.ptr(gvp), B, nz, nx, dt, h,
              _lib.stream_of(h1))
    return gvp, gh1, gh2


@acoustic_step_vjp.register_fake
def _(vp, d, h1, gy, dt, h):
    return torch.empty_like(vp), torch.empty_like(h1), torch.empty_like(h1)


def _acoustic_setup(ctx, inputs, output):
    vp, h1, h2, d, dt, h = inputs
    ctx.save_for_backward(vp, d, h1)
    ctx.dt, ctx.h = dt, h


def _acoustic_backward(ctx, gy):
    vp, d, h1 = ctx.saved_tensors
    gvp, gh1, gh2 = acoustic_step_vjp(vp, d, h1, gy, ctx.dt, ctx.h)
    return gvp, gh1, gh2, None, None, None


acoustic_step.register_autograd(_acoustic_backward, setup_context=_acoustic_setup)


# ------------------------------------------------------------- acoustic (HABC)

@torch.library.custom_op(f"{NS}::habc_step", mutates_args=())
def habc_step(vp: torch.Tensor, h1: torch.Tensor, h2: torch.Tensor, b: torch.Tensor, dt: float,
              h: float, multiple: bool) -> torch.Tensor:
    """Undamped step, then the HABC strips of b (seistorch/equations2d/acoustic_habc.py:206-221
    with habc() :147-204); multiple: free surface, no top strip."""
    h1, h2 = h1.contiguous(), h2.contiguous()
    _check_pair("habc_step", vp, h1, h2, b)
    B, nz, nx = h1.shape
    y = torch.empty_like(h1)
    _lib.call("seis_habc_step", _lib.ptr(vp), _lib.ptr(b), _lib.ptr(h1), _lib.ptr(h2),
              _lib.ptr(y), B, nz, nx, dt, h, int(multiple), _lib.stream_of(h1))
    return y


@habc_step.register_fake
def _(vp, h1, h2, b, dt, h, multiple):
    return torch.empty_like(h1, memory_format=torch.contiguous_format)


@torch.library.custom_op(f"{NS}::habc_step_vjp", mutates_args=())
def habc_step_vjp(vp: torch.Tensor, b: torch.Tensor, h1: torch.Tensor, h2: torch.Tensor,
                  gy: torch.Tensor, dt: float, h: float,
                  multiple: bool) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    h1, h2, gy = h1.contiguous(), h2.contiguous(), gy.contiguous()
    B, nz, nx = h1.shape
    gh1 = torch.empty_like(h1)
    gh2 = torch.empty_like(h1)
    gvp = torch.empty_like(vp)
    _lib.call("seis_habc_step_vjp", _lib.ptr(vp), _lib.ptr(b), _lib.ptr(h1), _lib.ptr(h2),
              _lib.ptr(gy), _lib.ptr(gh1), _lib.ptr(gh2), _lib.ptr(gvp), B, nz, nx, dt, h,
              int(multiple), _lib.stream_of(h1))
    return gvp, gh1, gh2


@habc_step_vjp.register_fake
def _(vp, b, h1, h2, gy, dt, h, multiple):
    return torch.empty_like(vp), torch.empty_like(h1), torch.empty_like(h1)


def _habc_setup(ctx, inputs, output):
    vp, h1, h2, b, dt, h, multiple = inputs
    ctx.save_for_backward(vp, b, h1, h2)
    ctx.meta = (dt, h, multiple)


def _habc_backward(ctx, gy):
    vp, b, h1, h2 = ctx.saved_tensors
    gvp, gh1, gh2 = habc_step_vjp(vp, b, h1, h2, gy, *ctx.meta)
    return gvp, gh1, gh2, None, None, None, None


habc_step.register_autograd(_habc_backward, setup_context=_habc_setup)


# ------------------------------------------------------------------- elastic

def elastic_coefficients(vp, vs, rho, d, dt, h):
    """(6, nz, nx): lam+2mu, lam, dt*mu/h, dt/(rho h), 1/(1+c), 1-c, with
    lam = rho(vp^2 - 2vs^2), mu = rho vs^2, c = dt*d/2 (elastic.py:11-14)."""
    _lib.require_device(vp, vs, rho, d)
    nz, nx = vp.shape
    for t in (vs, rho, d):
        if t.shape != vp.shape:
            raise RuntimeError("vp, vs, rho and d must share one (nz, nx) shape")
    cf = torch.empty((6, nz, nx), device=vp.device, dtype=torch.float32)
    _lib.call("seis_elastic_coefficients", _lib.ptr(vp), _lib.ptr(vs), _lib.ptr(rho),
              _lib.ptr(d), _lib.ptr(cf), nz, nx, dt, h, _lib.stream_of(vp))
    return cf


def elastic_shots_per_group(B):
    for s in (4, 2):
        if B >= s:
            return s
    return 1


_Fields = tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]


@torch.library.custom_op(f"{NS}::elastic_stage", mutates_args=())
def elastic_stage(vp: torch.Tensor, vs: torch.Tensor, rho: torch.Tensor, d: torch.Tensor,
                  vx: torch.Tensor, vz: torch.Tensor, txx: torch.Tensor, tzz: torch.Tensor,
                  txz: torch.Tensor, stage: int, sign: float, z0: int, z1: int, x0: int, x1: int,
                  dt: float, h: float) -> _Fields:
    """One stage of seistorch/equations2d/elastic.py:7-37 on the cells
    z0 <= z < z1, x0 <= x < x1 (stage 0: stresses from velocities, 1:
    velocities from stresses; sign -1: the reverse step of elastic.py:39-122),
    out of place: returns all five fields."""
    fin = [t.contiguous() for t in (vx, vz, txx, tzz, txz)]
    _lib.require_device(*fin)
    B, nz, nx = fin[0].shape
    for t in fin:
        if tuple(t.shape) != (B, nz, nx):
            raise RuntimeError("all five wavefields must share one (B, nz, nx) shape")
    if tuple(vp.shape) != (nz, nx):
        raise RuntimeError(f"model {tuple(vp.shape)} does not match fields {(nz, nx)}")
    cf = elastic_coefficients(vp, vs, rho, d, dt, h)
    out = [t.clone() for t in fin]
    _lib.call("seis_elastic_stage", stage, _lib.ptr(cf), _lib.ptr_array(out), sign, z0, z1, x0,
              x1, B, nz, nx, dt, h, elastic_shots_per_group(B), _lib.stream_of(vp))
    return tuple(out)


@elastic_stage.register_fake
def _(vp, vs, rho, d, vx, vz, txx, tzz, txz, stage, sign, z0, z1, x0, x1, dt, h):
    return tuple(torch.empty_like(vx, memory_format=torch.contiguous_format) for _ in range(5))


@torch.library.custom_op(f"{NS}::elastic_stage_vjp", mutates_args=())
def elastic_stage_vjp(vp: torch.Tensor, vs: torch.Tensor, rho: torch.Tensor, d: torch.Tensor,
                      vx: torch.Tensor, vz: torch.Tensor, txx: torch.Tensor, tzz: torch.Tensor,
                      txz: torch.Tensor, cvx: torch.Tensor, cvz: torch.Tensor,
                      ctxx: torch.Tensor, ctzz: torch.Tensor, ctxz: torch.Tensor, stage: int,
                      sign: float, z0: int, z1: int, x0: int, x1: int, dt: float,
                      h: float) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor,
                                         torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """(dJ/dvp, dJ/dvs, dJ/drho, cotangents of the five input fields) for the
    cotangents c* of the five output fields."""
    fin = [t.contiguous() for t in (vx, vz, txx, tzz, txz)]
    cin = [t.contiguous() for t in (cvx, cvz, ctxx, ctzz, ctxz)]
    B, nz, nx = fin[0].shape
    S = elastic_shots_per_group(B)
    G = (B + S - 1) // S
    cf = elastic_coefficients(vp, vs, rho, d, dt, h)
    cout = [torch.empty_like(fin[0]) for _ in range(5)]
    gco = torch.zeros((G, 4, nz, nx), device=vp.device, dtype=torch.float32)
    _lib.call("seis_elastic_stage_vjp", stage, _lib.ptr(cf), _lib.ptr_array(cin),
              _lib.ptr_array(cout), _lib.ptr_array(fin), _lib.ptr(gco), sign, z0, z1, x0, x1, B,
              nz, nx, dt, h, S, _lib.stream_of(vp))
    gvp, gvs, grho = (torch.empty_like(vp) for _ in range(3))
    _lib.call("seis_elastic_chain", _lib.ptr(gco), G, _lib.ptr(vp), _lib.ptr(vs), _lib.ptr(rho),
              _lib.ptr(gvp), _lib.ptr(gvs), _lib.ptr(grho), nz, nx, dt, h, _lib.stream_of(vp))
    return (gvp, gvs, grho, *cout)


@elastic_stage_vjp.register_fake
def _(vp, vs, rho, d, vx, vz, txx, tzz, txz, cvx, cvz, ctxx, ctzz, ctxz, stage, sign, z0, z1,
      x0, x1, dt, h):
    return (*(torch.empty_like(vp) for _ in range(3)),
            *(torch.empty_like(vx, memory_format=torch.contiguous_format) for _ in range(5)))


def _elastic_setup(ctx, inputs, output):
    vp, vs, rho, d, vx, vz, txx, tzz, txz = inputs[:9]
    ctx.save_for_backward(vp, vs, rho, d, vx, vz, txx, tzz, txz)
    ctx.meta = tuple(inputs[9:])


def _elastic_backward(ctx, *cots):
    saved = ctx.saved_tensors
    fin = saved[4:]
    cin = [c if c is not None else torch.zeros_like(fin[0]) for c in cots]
    g = elastic_stage_vjp(*saved, *cin, *ctx.meta)
    return (*g[:3], None, *g[3:], *([None] * len(ctx.meta)))


elastic_stage.register_autograd(_elastic_backward, setup_context=_elastic_setup)


# ------------------------------------------------------ variable density (HABC)

@torch.library.custom_op(f"{NS}::rho_habc_step", mutates_args=())
def rho_habc_step(vp: torch.Tensor, rho: torch.Tensor, p1: torch.Tensor, p2: torch.Tensor,
                  b: Optional[torch.Tensor], dt: float, h: float, multiple: bool,
                  blend: bool) -> torch.Tensor:
    """seistorch/equations2d/acoustic_rho_habc.py:31-57 (blend: the HABC
    strips of b applied) or its reverse step :59-89 before the ring restore
    (no blend, b unused)."""
    p1, p2 = p1.contiguous(), p2.contiguous()
    _lib.require_device(vp, rho, p1, p2, b)
    if p1.shape != p2.shape or p1.dim() != 3 or vp.shape != p1.shape[1:] \
            or rho.shape != vp.shape or (b is not None and b.shape != vp.shape):
        raise RuntimeError("rho_habc_step: shape mismatch")
    if blend and b is None:
        raise RuntimeError("rho_habc_step: the HABC blend needs the weight field b")
    B, nz, nx = p1.shape
    y = torch.empty_like(p1)
    _lib.call("seis_rho_habc_step", _lib.ptr(vp), _lib.ptr(rho), _lib.ptr(b), _lib.ptr(p1),
              _lib.ptr(p2), _lib.ptr(y), B, nz, nx, float(dt), float(h), int(multiple),
              int(blend), _lib.stream_of(p1))
    return y


@rho_habc_step.register_fake
def _(vp, rho, p1, p2, b, dt, h, multiple, blend):
    return torch.empty_like(p1, memory_format=torch.contiguous_format)
