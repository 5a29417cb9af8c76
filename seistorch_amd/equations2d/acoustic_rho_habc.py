"""Variable-density acoustic equation with the HABC boundary, backed by HIP
kernels.

Drop-in for ``seistorch.equations2d.acoustic_rho_habc``
(seistorch/equations2d/acoustic_rho_habc.py:31-121):

    _time_step(vp, rho, p1, p2, dt, h, b, habcs=masks)                -> (p', p1)
    _time_step_backward(vp, rho, p1, p2, dt, h, b, rings, source)     -> (p', p1)
    _time_step_backward_multiple(...)                                 -> (p', p1)

The update is p' = 2 p1 - p2 + dt^2 h^-2 rho vp^2 div(b grad p1) with the
buoyancy b = 1/rho averaged over the two cells of each face (Geiger and
Daley 2003, eq. 9), then the HABC strips of the weight field ``b`` (the
reference's argument name; ``seistorch_amd.habc.generate_habc_coefficients_2d``).
``propagate`` runs the whole loop (records) in one call; its backward is the
hand-written adjoint with boundary-saving reconstruction, which gives dJ/dvp
and dJ/drho on the undamped interior.  The per-step functions are forward
only: differentiating through them raises (the fused loop carries the
gradient).
"""
import torch

from .. import _lib
from .acoustic import _check_acquisition, check_finite
from .utils import restore_boundaries
from .acoustic_habc import _multiple


# shots sharing one thread of the forward (its per-cell set-up: face
# averages, strip weights and one-way coefficients, runs once for them)
DEFAULT_SHOTS_PER_THREAD = 16


def _no_gradient():
    raise NotImplementedError("the per-step variable-density functions are forward-only; "
                              "differentiate through propagate() (fused adjoint) instead")


def _step(vp, rho, p1, p2, b, dt, h, multiple, blend):
    _lib.require_device(vp, rho, p1, p2, b)
    if p1.shape != p2.shape or p1.dim() != 3 or vp.shape != p1.shape[1:] \
            or rho.shape != vp.shape or (b is not None and b.shape != vp.shape):
        raise RuntimeError("rho habc step: shape mismatch")
    B, nz, nx = p1.shape
    y = torch.empty_like(p1)
    _lib.call("seis_rho_habc_step", _lib.ptr(vp), _lib.ptr(rho), _lib.ptr(b),
              _lib.ptr(p1.contiguous()), _lib.ptr(p2.contiguous()), _lib.ptr(y), B, nz, nx,
              float(dt), float(h), int(multiple), int(blend), _lib.stream_of(p1))
    return y


class _RhoStepFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp, rho, p1, p2, b, dt, h, multiple, blend):
        return _step(vp, rho, p1, p2, b, dt, h, multiple, blend)

    @staticmethod
    def backward(ctx, gy):
        _no_gradient()


def _time_step(*args, **kwargs):
    """acoustic_rho_habc.py:31-57: (vp, rho, p1, p2, dt, h, b, habcs=) -> (p', p1)."""
    vp, rho = args[0:2]
    p1, p2 = args[2:4]
    dt, h, b = args[4:7]
    multiple = _multiple(kwargs.get("habcs"))
    return _RhoStepFn.apply(vp, rho, p1, p2, b, dt, h, multiple, True), p1


def _backward_common(args, multiple):
    vp, rho = args[0:2]
    p1, p2 = args[2:4]
    dt, h, _ = args[4:7]
    h_bd, _ = args[-2]
    src_type, src_func, src_values = args[-1]
    y = _RhoStepFn.apply(vp, rho, p1, p2, None, dt, h, multiple, False)
    with torch.no_grad():
        y = restore_boundaries(y, h_bd, multiple=multiple)
    y = src_func(y, src_values, 1)
    return y, p1


def _time_step_backward(*args, **kwargs):
    """acoustic_rho_habc.py:59-89: undamped reverse step, ring restore, source."""
    return _backward_common(args, multiple=False)


def _time_step_backward_multiple(*args, **kwargs):
    """acoustic_rho_habc.py:91-121."""
    return _backward_common(args, multiple=True)


class RhoHabcLoop(torch.autograd.Function):
    """The whole variable-density loop in one call; backward: the hand-written
    adjoint with boundary-saving reconstruction (ring at index 49, where the
    HABC weight is 0, plus periodic snapshots), giving dJ/dvp and dJ/drho on
    the undamped interior."""

    @staticmethod
    def forward(ctx, vp, rho, b, wavelet, src_zx, rec_flat, B, dt, h, multiple, S,
                snapshot_every=None, illum=None):
        from .acoustic import NPML, default_snapshot_every, ring_len
        _lib.require_device(vp, rho, b, wavelet, src_zx, rec_flat, illum)
        if illum is not None and illum.shape != vp.shape:
            raise RuntimeError("illum must have the model's shape")
        need_grad = ctx.needs_input_grad[0] or ctx.needs_input_grad[1]
        # the HABC weight is zero on the interior, which the adjoint's reversal needs
        _check_acquisition(vp, b, wavelet, src_zx, rec_flat, B, need_grad, multiple)
        if rho.shape != vp.shape:
            raise RuntimeError("rho must have the model's shape")
        nz, nx = vp.shape
        nt = wavelet.shape[0]
        R = rec_flat.shape[0]
        if S is None:
            S = min(B, DEFAULT_SHOTS_PER_THREAD)
        u = [torch.empty((B, nz, nx), device=vp.device, dtype=torch.float32) for _ in range(3)]
        rec = torch.empty((nt, R), device=vp.device, dtype=torch.float32)
        buoy = torch.empty_like(vp)
        ring = snaps = None
        K = 0
        if need_grad:
            if nt > 1:
                ring = torch.empty((nt - 1, B, ring_len(nz, nx, multiple=multiple)),
                                   device=vp.device)
            K = default_snapshot_every(nt, B, nz, nx, vp.device) if snapshot_every is None \
                else int(snapshot_every)
            n_snap = int(_lib.load().seis_snapshot_count(nt, K))
            if n_snap > 0:
                snaps = torch.empty((n_snap, 2, B, nz, nx), device=vp.device)
        _lib.call("seis_rho_habc_forward", _lib.ptr(vp), _lib.ptr(rho), _lib.ptr(b),
                  _lib.ptr(wavelet), nt, _lib.ptr(src_zx), _lib.ptr(rec_flat), R, _lib.ptr(rec),
                  _lib.ptr(ring), _lib.ptr(snaps), K if snaps is not None else 0,
                  _lib.ptr(buoy), _lib.ptr(illum), *[_lib.ptr(t) for t in u], B, nz, nx,
                  float(dt), float(h), NPML, int(multiple), int(S), _lib.stream_of(vp))
        check_finite(rec)
        if need_grad:
            ctx.save_for_backward(vp, rho, b, wavelet, src_zx, rec_flat)
            ctx.state = (ring, snaps, u[0], u[1], buoy)
            ctx.meta = (B, float(dt), float(h), int(S), K, int(multiple))
        return rec

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad_rec):
        from .acoustic import NPML
        vp, rho, b, wavelet, src_zx, rec_flat = ctx.saved_tensors
        ring, snaps, u_a, u_b, buoy = ctx.state
        ctx.state = None
        B, dt, h, S, K, multiple = ctx.meta
        nz, nx = vp.shape
        nt = wavelet.shape[0]
        R = rec_flat.shape[0]
        G = (B + S - 1) // S
        lam = [torch.empty_like(u_a) for _ in range(3)]
        gpart = torch.empty((2, G, nz, nx), device=vp.device)
        gvp = torch.empty_like(vp)
        grho = torch.empty_like(vp)
        _lib.call("seis_rho_habc_backward", _lib.ptr(vp), _lib.ptr(rho), _lib.ptr(b),
                  _lib.ptr(wavelet), nt, _lib.ptr(src_zx), _lib.ptr(rec_flat), R,
                  _lib.ptr(grad_rec.contiguous().float()), _lib.ptr(ring), _lib.ptr(snaps),
                  K if snaps is not None else 0, _lib.ptr(buoy), _lib.ptr(u_a), _lib.ptr(u_b),
                  *[_lib.ptr(t) for t in lam], _lib.ptr(gpart), _lib.ptr(gvp), _lib.ptr(grho), B,
                  nz, nx, dt, h, NPML, multiple, S, _lib.stream_of(vp))
        return (gvp, grho) + (None,) * 11


def propagate(vp, rho, b, wavelet, src_zx, rec_flat, B, dt, h, multiple=False,
              shots_per_group=None, illum=None, snapshot_every=None):
    """Records (nt, R) of B shots of the variable-density HABC loop,
    differentiable w.r.t. vp and rho (dJ/dvp and dJ/drho on the undamped
    interior: rows and columns [50, n-50), rows from 0 with a free surface).
    dt and h are taken as Python floats (double), as the reference's
    configuration scalars are.  illum: None, or an (nz, nx) float32 tensor
    filled with the source illumination (rnn.py:127-128, 204-205).
    snapshot_every: None (automatic), 0 (ring only, as the reference) or the
    interval of the forward snapshots that bound the rebuild's FP32 drift."""
    return RhoHabcLoop.apply(vp, rho, b, wavelet.contiguous(), src_zx.contiguous(),
                             rec_flat.contiguous(), int(B), float(dt), float(h), bool(multiple),
                             shots_per_group, snapshot_every, illum)
