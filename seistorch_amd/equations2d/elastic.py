"""2D isotropic elastic (velocity-stress, staggered) equation backed by HIP kernels.

Drop-in for ``seistorch.equations2d.elastic`` (seistorch/equations2d/elastic.py):

    _time_step(vp, vs, rho, vx, vz, txx, tzz, txz, dt, h, d)      -> 5 fields
    _time_step_backward(..., rings, [src_type, src_func, w])      -> 5 fields
    _time_step_backward_multiple(...)                             -> 5 fields

The step is the stress stage then the velocity stage (elastic.py:16-35); the
reference's reverse step (elastic.py:39-207) is the velocity stage then the
stress stage with negated time step on the interior, with ring restores and
the source re-added.  Every stage is the registered operator
torch.ops.seistorch_amd.elastic_stage, whose autograd formula is the
hand-written transpose kernel, so the autograd path (cell.py:74-75), the
reference's checkpoint mechanism (checkpoint.py) and its forced
torch.compile (model.py:83-85) all differentiate through HIP kernels.

``ElasticLoop`` / ``propagate`` is the fused path: the whole loop in one call,
backward = hand-written adjoint with boundary-saving reconstruction.
"""
import torch

from .. import _lib, library
from .acoustic import check_finite
from .utils import restore_boundaries

NPML = 49
N = 1
FIELDS = ("vx", "vz", "txx", "tzz", "txz")


def _f(x):
    return float(x.item()) if torch.is_tensor(x) else float(x)


def coefficients(vp, vs, rho, d, dt, h):
    """(6, nz, nx): lam+2mu, lam, dt*mu/h, dt/(rho h), 1/(1+c), 1-c."""
    return library.elastic_coefficients(vp, vs, rho, d, _f(dt), _f(h))


_shots = library.elastic_shots_per_group


def _stage(stage, sign, region, vp, vs, rho, d, fields, dt, h):
    """One stage as the registered operator torch.ops.seistorch_amd.elastic_stage
    (gradient: the hand-written transpose kernel and the chain to vp, vs, rho)."""
    z0, z1, x0, x1 = (int(v) for v in region)
    return library.elastic_stage(vp, vs, rho, d, *fields, int(stage), float(sign), z0, z1, x0,
                                 x1, _f(dt), _f(h))


def _time_step(*args, **kwargs):
    """elastic.py:7-37: stresses then velocities, returns (vx, vz, txx, tzz, txz)."""
    vp, vs, rho = args[0:3]
    fields = args[3:8]
    dt, h, d = args[8:11]
    nz, nx = vp.shape[-2:]
    full = (0, nz, 0, nx)
    out = _stage(0, 1.0, full, vp, vs, rho, d, fields, dt, h)
    return _stage(1, 1.0, full, vp, vs, rho, d, out, dt, h)


def _backward_common(args, top):
    vp, vs, rho = args[0:3]
    fields = args[3:8]
    dt, h, d = args[8:11]
    rings = args[-2]
    src_type, src_func, src_values = args[-1]
    nz, nx = vp.shape[-2:]
    region = (top if top is not None else NPML + N, nz - NPML - N, NPML + N, nx - NPML - N)
    multiple = top is not None
    f = list(_stage(1, -1.0, region, vp, vs, rho, d, fields, dt, h))
    for k in (0, 1):
        f[k] = restore_boundaries(f[k], rings[k], multiple=multiple)
    f = list(_stage(0, -1.0, region, vp, vs, rho, d, f, dt, h))
    for k in (2, 3, 4):
        f[k] = restore_boundaries(f[k], rings[k], multiple=multiple)
    for s_type in src_type:
        k = FIELDS.index(s_type)
        f[k].data.copy_(src_func(f[k], src_values, -1))
    return tuple(f)


def _time_step_backward(*args, **kwargs):
    """elastic.py:39-122: reverse step on the interior, rings restored, source re-added."""
    rings = list(args[-2])
    return _backward_common(args[:-2] + (rings, args[-1]), top=None)


def _time_step_backward_multiple(*args, **kwargs):
    """elastic.py:124-207: as above with the top rows included (free surface)."""
    rings = list(args[-2])
    return _backward_common(args[:-2] + (rings, args[-1]), top=0)


# ----------------------------------------------------------------- fused loop

SNAP_MIN_EVERY = 50
SNAP_HBM_FRACTION = 0.25


def default_snapshot_every(nt, B, nz, nx, device):
    state = 5 * B * nz * nx * 4
    total = torch.cuda.get_device_properties(device).total_memory
    max_snaps = int(SNAP_HBM_FRACTION * total // state)
    if max_snaps < 1 or nt <= SNAP_MIN_EVERY:
        return 0
    return max(SNAP_MIN_EVERY, -(-(nt - 1) // max_snaps))


def _mask(names):
    m = 0
    for n in names:
        if n not in FIELDS:
            raise RuntimeError(f"unknown elastic field {n!r}; expected one of {FIELDS}")
        m |= 1 << FIELDS.index(n)
    return m


class ElasticLoop(torch.autograd.Function):
    """Whole elastic loop -> records (nt, R, nch); backward -> (dvp, dvs, drho)."""

    @staticmethod
    def forward(ctx, vp, vs, rho, d, wavelet, src_zx, rec_flat, B, dt, h, src_type, rec_type,
                shots_per_group, snapshot_every, fused=True, illum=None, multiple=False):
        _lib.require_device(vp, vs, rho, d, wavelet, src_zx, rec_flat, illum)
        if illum is not None and illum.shape != vp.shape:
            raise RuntimeError("illum must have the model's shape")
        from .acoustic import _check_acquisition
        need_grad = any(ctx.needs_input_grad[:3])
        _check_acquisition(vp, d, wavelet, src_zx, rec_flat, B, need_grad, multiple)
        nz, nx = vp.shape
        nt = wavelet.shape[0]
        R = rec_flat.shape[0]
        if shots_per_group is None:
            from .. import autotune
            shots_per_group = autotune.lookup("elastic", B, vp.shape[0], vp.shape[1])
        if isinstance(shots_per_group, (tuple, list)):
            S, S_bwd = (int(v) for v in shots_per_group)
        else:
            S = S_bwd = int(shots_per_group or _shots(B))
        cf = coefficients(vp, vs, rho, d, dt, h)
        fields = [torch.empty((B, nz, nx), device=vp.device) for _ in range(5)]
        rf = torch.tensor([FIELDS.index(n) for n in rec_type], dtype=torch.int32,
                          device=vp.device)
        nch = len(rec_type)
        rec = torch.empty((nt, R, nch), device=vp.device)
        ring = snaps = None
        K = 0
        if need_grad:
            L = int(_lib.load().seis_ring_len(nz, nx, NPML, int(multiple)))
            if nt > 1:
                ring = torch.empty((nt - 1, 5, B, L), device=vp.device)
            K = default_snapshot_every(nt, B, nz, nx, vp.device) if snapshot_every is None \
                else int(snapshot_every)
            n_snap = (nt - 1) // K if K > 0 else 0
            if n_snap > 0:
                snaps = torch.empty((n_snap, 5, B, nz, nx), device=vp.device)
        scratch = [torch.empty_like(fields[0]) for _ in range(5)] if fused else None
        _lib.call("seis_elastic_forward", _lib.ptr(cf), _lib.ptr(wavelet), nt, _lib.ptr(src_zx),
                  _mask(src_type), _lib.ptr(rec_flat), R, _lib.ptr(rf), nch, _lib.ptr(rec),
                  _lib.ptr(ring), _lib.ptr(snaps), K, _lib.ptr(illum),
                  FIELDS.index(src_type[-1]), _lib.ptr_array(fields),
                  _lib.ptr_array(scratch) if fused else None, B, nz, nx, dt, h, NPML,
                  int(multiple), S, _lib.stream_of(vp))
        del scratch  # ping-pong buffers of the fused steps
        check_finite(rec)
        if need_grad:
            ctx.save_for_backward(vp, vs, rho, cf, wavelet, src_zx, rec_flat, rf)
            ctx.state = (ring, snaps, fields)
            ctx.meta = (B, dt, h, S_bwd, K, _mask(src_type), nch, bool(fused), int(multiple))
        return rec

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad_rec):
        vp, vs, rho, cf, wavelet, src_zx, rec_flat, rf = ctx.saved_tensors
        ring, snaps, fields = ctx.state
        ctx.state = None
        B, dt, h, S, K, smask, nch, fused, multiple = ctx.meta
        nz, nx = vp.shape
        nt = wavelet.shape[0]
        R = rec_flat.shape[0]
        G = (B + S - 1) // S
        grad_rec = grad_rec.contiguous().to(torch.float32)
        adj_a = [torch.empty_like(fields[0]) for _ in range(5)]
        adj_b = [torch.empty_like(fields[0]) for _ in range(5)]
        scratch = [torch.empty_like(fields[0]) for _ in range(5)] if fused else None
        gco = torch.empty((G, 4, nz, nx), device=vp.device)
        _lib.call("seis_elastic_backward", _lib.ptr(cf), _lib.ptr(wavelet), nt,
                  _lib.ptr(src_zx), smask, _lib.ptr(rec_flat), R, _lib.ptr(rf), nch,
                  _lib.ptr(grad_rec), _lib.ptr(ring), _lib.ptr(snaps),
                  K if snaps is not None else 0, _lib.ptr_array(fields), _lib.ptr_array(adj_a),
                  _lib.ptr_array(adj_b), _lib.ptr_array(scratch) if fused else None,
                  _lib.ptr(gco), B, nz, nx, dt, h, NPML, multiple, S, _lib.stream_of(vp))
        gvp, gvs, grho = (torch.empty_like(vp) for _ in range(3))
        _lib.call("seis_elastic_chain", _lib.ptr(gco), G, _lib.ptr(vp), _lib.ptr(vs),
                  _lib.ptr(rho), _lib.ptr(gvp), _lib.ptr(gvs), _lib.ptr(grho), nz, nx, dt, h,
                  _lib.stream_of(vp))
        return (gvp, gvs, grho) + (None,) * 14


def propagate(vp, vs, rho, d, wavelet, src_zx, rec_flat, B, dt, h, src_type=("txx", "tzz"),
              rec_type=("vx", "vz"), shots_per_group=None, snapshot_every=None, fused=True,
              illum=None, multiple=False):
    """Records (nt, R, nch) of B shots; differentiable w.r.t. vp, vs, rho.
    shots_per_group: shots per thread (1, 2, 4, 8), or a pair (forward,
    adjoint); None: the tuned value for this shape, else a default.
    fused: each forward step and each reverse step of the adjoint in one pass
    (five more field buffers).  illum: None, or an (nz, nx) float32 tensor
    filled with the source illumination of the last source type's field
    (rnn.py:204-205).  multiple: free surface (d from
    ``pml.generate_pml_coefficients_2d(..., multiple=True)``); the gradients
    then cover rows [0, nz-50) of the interior columns, the region of
    elastic.py:124-207.  Non-finite records raise ValueError (rnn.py:215)."""
    return ElasticLoop.apply(vp, vs, rho, d, wavelet.contiguous(), src_zx.contiguous(),
                             rec_flat.contiguous(), int(B), _f(dt), _f(h), tuple(src_type),
                             tuple(rec_type), shots_per_group, snapshot_every, bool(fused),
                             illum, bool(multiple))
