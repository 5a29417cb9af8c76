"""Constant-density acoustic equation with PML, backed by HIP kernels.

Drop-in for ``seistorch.equations2d.acoustic`` (seistorch/equations2d/acoustic.py):
the same three functions with the same positional arguments,

    _time_step(vp, h1, h2, dt, h, d, habcs=None)              -> (y, h1)
    _time_step_backward(vp, h1, h2, dt, h, d, rings, source)   -> (y, h1)
    _time_step_backward_multiple(...)                          -> (y, h1)

so ``equation: acoustic`` in a YAML config selects it.  The per-step
functions are differentiable (the reference's ``boundary_saving: false`` path
backpropagates through them, seistorch/cell.py:74-75); their gradient is the
hand-written VJP kernel, not autograd of elementwise ops.  They call the
``torch.library`` operators of ``seistorch_amd.library``, so the reference's
forced ``torch.compile`` of both functions (seistorch/model.py:83-85) traces
them without breaking.

``AcousticLoop`` is the fused alternative used by ``WaveRNN`` when the
configuration allows it: the whole time loop (step, source, receivers, ring
saving) in one library call, and its backward the hand-written adjoint with
boundary-saving reconstruction, replacing autograd-through-the-loop
(seistorch/checkpoint_new.py:109-217).
"""
import torch

from .. import _lib, library
from .utils import restore_boundaries

NPML = 49  # ring index of the boundary saving (seistorch/equations2d/utils.py:50)


def _f(x):
    return float(x.item()) if torch.is_tensor(x) else float(x)


def step(vp, h1, h2, d, dt, h):
    """y = A*(2/dt^2*h1 - B*h2 + vp^2*L h1) (acoustic.py:81-84): the registered
    operator torch.ops.seistorch_amd.acoustic_step, whose gradient is the
    hand-written VJP kernel."""
    return library.acoustic_step(vp, h1, h2, d, _f(dt), _f(h))


def _time_step(*args, **kwargs):
    """acoustic.py:73-86: returns (y, h1)."""
    c = args[0]
    h1, h2 = args[1:3]
    dt, h, b = args[3:6]
    return step(c, h1, h2, b, dt, h), h1


def _backward_common(args, multiple):
    vp = args[0]
    h1, h2 = args[1:3]
    dt, h, b = args[3:6]
    h_bd, _ = args[-2]
    src_type, src_func, src_values = args[-1]
    y = step(vp, h1, h2, b, dt, h)
    with torch.no_grad():
        y = restore_boundaries(y, h_bd, multiple=multiple)
    y = src_func(y, src_values, 1)
    return y, h1


def _time_step_backward(*args, **kwargs):
    """acoustic.py:88-112: reverse step, ring restore, source re-add."""
    return _backward_common(args, multiple=False)


def _time_step_backward_multiple(*args, **kwargs):
    """acoustic.py:114-137: as above with the top ring left alone."""
    return _backward_common(args, multiple=True)


# ----------------------------------------------------------------- fused loop

def ring_len(nz, nx, npml=NPML, multiple=False):
    """Cells of the boundary-saving ring; multiple: free surface, no top row."""
    return int(_lib.load().seis_ring_len(nz, nx, npml, int(bool(multiple))))


def check_finite(rec):
    """The reference's loop raises ValueError on NaN records (rnn.py:215 with
    type.py:41-46); a time step beyond the stability limit shows up here."""
    if not bool(torch.isfinite(rec).all()):
        raise ValueError("the records contain non-finite values (is dt within the CFL limit?)")
    return rec


SNAP_MIN_EVERY = 50        # densest snapshot spacing chosen automatically
SNAP_HBM_FRACTION = 0.25   # share of device memory the snapshots may take


def default_snapshot_every(nt, B, nz, nx, device):
    """Snapshot spacing: as dense as SNAP_MIN_EVERY allows within the budget.

    The pair (u^m, u^{m+1}) every K steps bounds the FP32 drift of the
    reverse-time reconstruction to K steps; 0 means pure boundary saving.
    """
    pair = 2 * B * nz * nx * 4
    total = torch.cuda.get_device_properties(device).total_memory
    max_snaps = int(SNAP_HBM_FRACTION * total // pair)
    if max_snaps < 1 or nt <= SNAP_MIN_EVERY + 1:
        return 0
    return max(SNAP_MIN_EVERY, -(-(nt - 2) // max_snaps))


# launch shape when no tuned value is known: rows per thread (cells per lane
# along x is the widest of 4, 2, 1 that divides nx)
ROWS_PER_THREAD = 2


def default_vec_x(nx):
    return 4 if nx % 4 == 0 else 2 if nx % 2 == 0 else 1


def default_shots_per_group(B):
    """Shots advanced by one thread: amortises model and gradient traffic."""
    for s in (4, 2):
        if B >= s:
            return s
    return 1


def _launch_params(B, nz, nx, shape):
    """(shots_per_group, rows_per_thread, vec_x): caller's, tuned, or default."""
    from .. import autotune
    tuned = autotune.lookup("acoustic", B, nz, nx)
    default = tuned if tuned else (default_shots_per_group(B), ROWS_PER_THREAD, default_vec_x(nx))
    shape = tuple(shape) if shape is not None else (None, None, None)
    return tuple(int(v if v is not None else dv) for v, dv in zip(shape, default))


def _check_acquisition(vp, d, wavelet, src_zx, rec_flat, B, need_grad, multiple=False):
    """Host-side bounds checks: the kernels index with these values unchecked.
    With a gradient the damping must vanish on the undamped interior that the
    adjoint rebuilds: rows [50, nz-50), or [0, nz-50) with a free surface
    (multiple), and columns [50, nx-50)."""
    nz, nx = vp.shape
    if vp.dim() != 2 or d.shape != vp.shape:
        raise RuntimeError(f"vp {tuple(vp.shape)} and d {tuple(d.shape)} must be equal (nz, nx)")
    if wavelet.dim() != 1 or wavelet.shape[0] < 1 or wavelet.dtype != torch.float32:
        raise RuntimeError("wavelet must be a non-empty float32 vector (nt,)")
    if src_zx.dtype != torch.int32 or tuple(src_zx.shape) != (B, 2):
        raise RuntimeError(f"src_zx must be int32 ({B}, 2), got {src_zx.dtype} {tuple(src_zx.shape)}")
    if rec_flat.dtype != torch.int64 or rec_flat.dim() != 1:
        raise RuntimeError("rec_flat must be an int64 vector")
    lo = torch.stack([src_zx.min(), src_zx[:, 0].max() - nz, src_zx[:, 1].max() - nx])
    bad_src = bool((lo[0] < 0) | (lo[1] >= 0) | (lo[2] >= 0))
    bad_rec = rec_flat.numel() > 0 and bool((rec_flat.min() < 0) | (rec_flat.max() >= B * nz * nx))
    if bad_src:
        raise RuntimeError(f"a source lies outside the {nz}x{nx} grid")
    if bad_rec:
        raise RuntimeError(f"a receiver index lies outside the ({B}, {nz}, {nx}) fields")
    if need_grad:
        m = NPML + 1
        z0 = 0 if multiple else m
        if nz - m <= z0 or nx <= 2 * m:
            raise RuntimeError(f"grid {nz}x{nx} has no interior inside the 50-cell boundary")
        if bool(d[z0:nz - m, m:nx - m].abs().max() != 0):
            raise RuntimeError("boundary-saving adjoint needs d == 0 inside the 50-cell boundary"
                               + (" (free surface: from row 0)" if multiple else ""))


def shared_receiver_map(rec_flat, B, nz, nx):
    """(ok, map): when every shot has its receivers at the same cells in the
    same order (rec_flat[b*Rs + r] = b*nz*nx + cell_r, no cell twice), map is
    the (nz, nx) int32 receiver index of each cell (-1 elsewhere), which the
    two-step adjoint needs; ok is False otherwise.  R == 0 gives (True, None)."""
    R = rec_flat.numel()
    if R == 0:
        return True, None
    if R % B:
        return False, None
    Rs = R // B
    plane = nz * nx
    rf = rec_flat.view(B, Rs)
    cells = rf[0]
    shift = torch.arange(B, device=rec_flat.device, dtype=torch.int64)[:, None] * plane
    if not torch.equal(rf - shift, cells.expand(B, Rs)):
        return False, None
    if bool((cells < 0).any()) or bool((cells >= plane).any()):
        return False, None
    m = torch.full((plane,), -1, dtype=torch.int32, device=rec_flat.device)
    m[cells] = torch.arange(Rs, dtype=torch.int32, device=rec_flat.device)
    if int((m >= 0).sum()) != Rs:  # a cell holds two receivers
        return False, None
    return True, m.view(nz, nx)


class AcousticLoop(torch.autograd.Function):
    """Whole forward loop -> records (nt, R); backward = hand-written adjoint.

    Inputs: vp, d (nz, nx); wavelet (nt,); src_zx (B, 2) int32 padded (z, x);
    rec_flat (R,) int64 flat indices into (B, nz, nx); B; dt; h; launch shape
    (shots_per_group, rows_per_thread, vec_x), None entries automatic;
    snapshot_every (None: automatic, 0: ring only as in the reference);
    two_step: two time steps per forward launch (two more field buffers);
    two_step_adjoint: two reverse steps per adjoint launch (four more, when
    every shot records at the same receiver cells; slower than one step per
    launch at present, kept for its bit-identical results); illum: None, or
    an (nz, nx) float32 tensor that receives the source illumination (the sum
    over steps and shots of the squared post-source field, rnn.py:204-205).
    """

    @staticmethod
    def forward(ctx, vp, d, wavelet, src_zx, rec_flat, B, dt, h, shape, snapshot_every,
                two_step=True, two_step_adjoint=False, illum=None, multiple=False,
                extra_src=None, extra_wav=None):
        _lib.require_device(vp, d, wavelet, src_zx, rec_flat, illum, extra_src, extra_wav)
        n_extra = 0 if extra_src is None else int(extra_src.shape[0])
        if n_extra:
            two_step = False  # the extra sources are added between single steps
        nz, nx = vp.shape
        if illum is not None and (illum.shape != (nz, nx) or illum.dtype != torch.float32
                                  or not illum.is_contiguous()):
            raise RuntimeError(f"illum must be a contiguous float32 ({nz}, {nx}) tensor")
        nt = wavelet.shape[0]
        R = rec_flat.shape[0]
        S, rz, vx = _launch_params(B, nz, nx, shape)
        need_grad = ctx.needs_input_grad[0]
        _check_acquisition(vp, d, wavelet, src_zx, rec_flat, B, need_grad, multiple)
        u_a = torch.empty((B, nz, nx), device=vp.device, dtype=torch.float32)
        u_b = torch.empty_like(u_a)
        rec = torch.empty((nt, R), device=vp.device, dtype=torch.float32)
        ring = None
        if need_grad and nt > 1:
            L = ring_len(nz, nx, multiple=multiple)
            if L <= 0:
                raise RuntimeError(f"grid {nz}x{nx} has no interior inside the 50-cell boundary")
            ring = torch.empty((nt - 1, B, L), device=vp.device, dtype=torch.float32)
        K = 0
        snaps = None
        if need_grad:
            K = default_snapshot_every(nt, B, nz, nx, vp.device) if snapshot_every is None \
                else int(snapshot_every)
            n_snap = int(_lib.load().seis_snapshot_count(nt, K))
            if n_snap > 0:
                snaps = torch.empty((n_snap, 2, B, nz, nx), device=vp.device, dtype=torch.float32)
        u_c = u_d = None
        if two_step and nt >= 2:
            u_c = torch.empty_like(u_a)
            u_d = torch.empty_like(u_a)
        # with two steps per launch the launches sum the illumination per shot group
        ipart = None
        if illum is not None and u_c is not None:
            ipart = torch.empty(((B + S - 1) // S, nz, nx), device=vp.device, dtype=torch.float32)
        _lib.call("seis_acoustic_forward", _lib.ptr(vp), _lib.ptr(d), _lib.ptr(wavelet), nt,
                  _lib.ptr(src_zx), _lib.ptr(extra_src), _lib.ptr(extra_wav), n_extra,
                  _lib.ptr(rec_flat), R, _lib.ptr(rec), _lib.ptr(ring),
                  _lib.ptr(snaps), K, _lib.ptr(illum), _lib.ptr(ipart), _lib.ptr(u_a),
                  _lib.ptr(u_b),
                  _lib.ptr(u_c), _lib.ptr(u_d), B, nz, nx, dt, h, NPML, int(multiple), S, rz,
                  vx, _lib.stream_of(vp))
        del u_c, u_d  # scratch of the two-step launches
        check_finite(rec)
        if need_grad:
            ctx.save_for_backward(vp, d, wavelet, src_zx, rec_flat)
            ctx.state = (ring, snaps, u_a, u_b, extra_src, extra_wav)
            ctx.meta = (B, dt, h, S, K, rz, vx, bool(two_step_adjoint) and not n_extra,
                        int(multiple))
        return rec

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad_rec):
        vp, d, wavelet, src_zx, rec_flat = ctx.saved_tensors
        ring, snaps, u_a, u_b, extra_src, extra_wav = ctx.state
        ctx.state = None  # the backward consumes the final forward state
        n_extra = 0 if extra_src is None else int(extra_src.shape[0])
        B, dt, h, S, K, rz, vx, two_step_adjoint, multiple = ctx.meta
        nz, nx = vp.shape
        nt = wavelet.shape[0]
        R = rec_flat.shape[0]
        G = (B + S - 1) // S
        grad_rec = grad_rec.contiguous().to(torch.float32)
        mu_a = torch.empty_like(u_a)
        mu_b = torch.empty_like(u_a)
        gpart = torch.empty((G, nz, nx), device=vp.device, dtype=torch.float32)
        grad = torch.empty_like(vp)
        extra = [None] * 5  # u_c, u_d, mu_c, mu_d, rec_map of the two-step adjoint
        if two_step_adjoint and nt >= 4 and ring is not None:
            ok, rmap = shared_receiver_map(rec_flat, B, nz, nx)
            if ok:
                extra = [torch.empty_like(u_a) for _ in range(4)] + [rmap]
        _lib.call("seis_acoustic_backward", _lib.ptr(vp), _lib.ptr(d), _lib.ptr(wavelet), nt,
                  _lib.ptr(src_zx), _lib.ptr(extra_src), _lib.ptr(extra_wav), n_extra,
                  _lib.ptr(rec_flat), R, _lib.ptr(grad_rec), _lib.ptr(ring),
                  _lib.ptr(snaps), K if snaps is not None else 0, _lib.ptr(u_a), _lib.ptr(u_b),
                  _lib.ptr(mu_a), _lib.ptr(mu_b), *[_lib.ptr(t) for t in extra],
                  _lib.ptr(gpart), _lib.ptr(grad), B, nz, nx, dt, h, NPML, multiple, S, rz, vx,
                  _lib.stream_of(vp))
        return (grad,) + (None,) * 15


def _check_extra_sources(extra_sources, B, nz, nx, nt, need_grad, multiple):
    """(extra_src int32 (n, 3) of (shot, z, x), extra_wav float32 (n, nt)) or
    (None, None); the kernels index with these values unchecked."""
    if extra_sources is None:
        return None, None
    es, ew = extra_sources
    if es.dtype != torch.int32 or es.dim() != 2 or es.shape[1] != 3:
        raise RuntimeError("extra source table must be int32 (n, 3): shot, z, x")
    if ew.dtype != torch.float32 or tuple(ew.shape) != (es.shape[0], nt):
        raise RuntimeError(f"extra wavelets must be float32 ({es.shape[0]}, {nt})")
    if es.shape[0] == 0:
        return None, None
    lo = 0 if multiple else NPML + 1
    zlo, zhi, xlo, xhi = (lo, nz - NPML - 1, NPML + 1, nx - NPML - 1) if need_grad \
        else (0, nz, 0, nx)
    bad = (es[:, 0] < 0) | (es[:, 0] >= B) | (es[:, 1] < zlo) | (es[:, 1] >= zhi) \
        | (es[:, 2] < xlo) | (es[:, 2] >= xhi)
    if bool(bad.any()):
        raise RuntimeError("an extra source lies outside the grid"
                           + (" interior that the adjoint rebuilds" if need_grad else ""))
    return es.contiguous(), ew.contiguous()


def propagate(vp, d, wavelet, src_zx, rec_flat, B, dt, h, shots_per_group=None,
              snapshot_every=None, rows_per_thread=None, vec_x=None, two_step=True,
              two_step_adjoint=False, illum=None, multiple=False, extra_sources=None):
    """Records (nt, R) of B shots; differentiable w.r.t. vp through the adjoint.

    The launch shape -- shots_per_group (1, 2, 4, 8), rows_per_thread
    (1, 2, 4), vec_x (cells per lane along x: 1, 2, 4, dividing nx) --
    defaults to the autotuned values for this shape (``seistorch_amd.autotune``),
    else to fixed defaults.  two_step runs two forward time steps per launch
    (same results, half the forward's field traffic).  two_step_adjoint runs
    two reverse steps per adjoint launch (same results up to the order of the
    gradient additions at receiver cells; at present slower than one step per
    launch: its fields of two iterations need ~250 registers per lane).
    illum: None, or an (nz, nx) float32 tensor filled with the source
    illumination of this forward (rnn.py:127-128, 204-205), the divisor of
    ``process.PostProcess.precondition``.
    multiple: free surface (the reference's geom.multiple): d damps no top
    strip (``pml.generate_pml_coefficients_2d(..., multiple=True)``), the ring
    has no top row and the gradient covers rows [0, nz-50) of the interior
    columns, as the reference's _time_step_backward_multiple rebuilds them
    (acoustic.py:114-137).  Non-finite records raise ValueError (rnn.py:215).
    extra_sources: None, or (table, wavelets): more point sources, each with
    its own wavelet, added after every step like the shot's own source (the
    source encoding of codingfwi.py:184-334, source.py:54-55: many encoded
    sources in one field); table int32 (n, 3) rows (shot, z, x), wavelets
    float32 (n, nt).  With a gradient they must lie in the rebuilt interior.
    They run one time step per launch.
    """
    nt = wavelet.shape[0]
    es, ew = _check_extra_sources(extra_sources, int(B), vp.shape[0], vp.shape[1], nt,
                                  vp.requires_grad and torch.is_grad_enabled(), multiple)
    return AcousticLoop.apply(vp, d, wavelet.contiguous(), src_zx.contiguous(),
                              rec_flat.contiguous(), int(B), _f(dt), _f(h),
                              (shots_per_group, rows_per_thread, vec_x), snapshot_every,
                              bool(two_step), bool(two_step_adjoint), illum, bool(multiple), es,
                              ew)
