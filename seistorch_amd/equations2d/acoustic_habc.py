"""Constant-density acoustic equation with the HABC boundary, backed by HIP kernels.

Drop-in for ``seistorch.equations2d.acoustic_habc``
(seistorch/equations2d/acoustic_habc.py):

    _time_step(vp, h1, h2, dt, h, b, habcs=masks)             -> (y, h1)
    _time_step_backward(vp, h1, h2, dt, h, b, rings, source)   -> (y, h1)
    _time_step_backward_multiple(...)                          -> (y, h1)

``b`` is the HABC weight field (seistorch_amd.habc.generate_habc_coefficients_2d);
``habcs`` are the strip masks the reference's cell passes (cell.py:20-27); only
whether the top mask is None (free surface) is read from them, the kernel
evaluates the trapezoids itself.  The step is the registered operator
torch.ops.seistorch_amd.habc_step; its gradient is the hand-written transpose
kernel.  ``HabcLoop`` / ``propagate`` is the fused loop with the
boundary-saving adjoint (ring at index 49, whose HABC weight is 0, so the
interior reverses exactly as in the PML case).
"""
import torch

from .. import _lib, library
from .acoustic import (NPML, _check_acquisition, _f, check_finite, default_shots_per_group,
                       restore_boundaries, step as _undamped_step)


def _multiple(habcs):
    return habcs is not None and habcs[0] is None


def _time_step(*args, **kwargs):
    """acoustic_habc.py:206-221: undamped step + HABC strips, returns (y, h1)."""
    c = args[0]
    h1, h2 = args[1:3]
    dt, h, b = args[3:6]
    multiple = _multiple(kwargs.get("habcs"))
    y = library.habc_step(c, h1, h2, b, _f(dt), _f(h), multiple)
    return y, h1


def _backward_common(args, multiple):
    vp = args[0]
    h1, h2 = args[1:3]
    dt, h, b = args[3:6]
    h_bd, _ = args[-2]
    src_type, src_func, src_values = args[-1]
    y = _undamped_step(vp, h1, h2, torch.zeros_like(vp), dt, h)
    with torch.no_grad():
        y = restore_boundaries(y, h_bd, multiple=multiple)
    y = src_func(y, src_values, 1)
    return y, h1


def _time_step_backward(*args, **kwargs):
    """acoustic_habc.py:223-245: undamped reverse step, ring restore, source re-add."""
    return _backward_common(args, multiple=False)


def _time_step_backward_multiple(*args, **kwargs):
    """acoustic_habc.py:247-268."""
    return _backward_common(args, multiple=True)


# ----------------------------------------------------------------- fused loop

class HabcLoop(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp, b, wavelet, src_zx, rec_flat, B, dt, h, multiple, shots_per_group,
                snapshot_every, illum=None):
        from .acoustic import default_snapshot_every, ring_len
        _lib.require_device(vp, b, wavelet, src_zx, rec_flat, illum)
        if illum is not None and illum.shape != vp.shape:
            raise RuntimeError("illum must have the model's shape")
        need_grad = ctx.needs_input_grad[0]
        # the HABC weight is zero on the interior, which the adjoint's reversal needs
        _check_acquisition(vp, b, wavelet, src_zx, rec_flat, B, need_grad, multiple)
        nz, nx = vp.shape
        nt = wavelet.shape[0]
        R = rec_flat.shape[0]
        if shots_per_group is None:
            from .. import autotune
            tuned = autotune.lookup("habc", B, nz, nx)
            shots_per_group = tuned[0] if tuned else None
        S = int(shots_per_group or default_shots_per_group(B))
        u = [torch.empty((B, nz, nx), device=vp.device) for _ in range(3)]
        rec = torch.empty((nt, R), device=vp.device)
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
        _lib.call("seis_habc_forward", _lib.ptr(vp), _lib.ptr(b), _lib.ptr(wavelet), nt,
                  _lib.ptr(src_zx), _lib.ptr(rec_flat), R, _lib.ptr(rec), _lib.ptr(ring),
                  _lib.ptr(snaps), K, _lib.ptr(illum), _lib.ptr(u[0]), _lib.ptr(u[1]),
                  _lib.ptr(u[2]), B, nz, nx, dt, h, NPML, int(multiple), S, _lib.stream_of(vp))
        check_finite(rec)
        if need_grad:
            ctx.save_for_backward(vp, b, wavelet, src_zx, rec_flat)
            ctx.state = (ring, snaps, u[0], u[1])
            ctx.meta = (B, dt, h, S, K, int(multiple))
        return rec

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad_rec):
        vp, b, wavelet, src_zx, rec_flat = ctx.saved_tensors
        ring, snaps, u_a, u_b = ctx.state
        ctx.state = None
        B, dt, h, S, K, multiple = ctx.meta
        nz, nx = vp.shape
        nt = wavelet.shape[0]
        R = rec_flat.shape[0]
        G = (B + S - 1) // S
        lam = [torch.empty_like(u_a) for _ in range(3)]
        gpart = torch.empty((G, nz, nx), device=vp.device)
        grad = torch.empty_like(vp)
        _lib.call("seis_habc_backward", _lib.ptr(vp), _lib.ptr(b), _lib.ptr(wavelet), nt,
                  _lib.ptr(src_zx), _lib.ptr(rec_flat), R,
                  _lib.ptr(grad_rec.contiguous().float()), _lib.ptr(ring), _lib.ptr(snaps),
                  K if snaps is not None else 0, _lib.ptr(u_a), _lib.ptr(u_b), _lib.ptr(lam[0]),
                  _lib.ptr(lam[1]), _lib.ptr(lam[2]), _lib.ptr(gpart), _lib.ptr(grad), B, nz, nx,
                  dt, h, NPML, multiple, S, _lib.stream_of(vp))
        return (grad,) + (None,) * 11


def propagate(vp, b, wavelet, src_zx, rec_flat, B, dt, h, multiple=False, shots_per_group=None,
              snapshot_every=None, illum=None):
    """Records (nt, R) of the HABC loop; differentiable w.r.t. vp.  illum:
    None, or an (nz, nx) float32 tensor filled with the source illumination
    (rnn.py:127-128, 204-205).  multiple: free surface (no top strip); the
    gradient then covers rows [0, nz-50) of the interior columns
    (acoustic_habc.py:247-267).  Non-finite records raise ValueError."""
    return HabcLoop.apply(vp, b, wavelet.contiguous(), src_zx.contiguous(), rec_flat.contiguous(),
                          int(B), _f(dt), _f(h), bool(multiple), shots_per_group, snapshot_every,
                          illum)
