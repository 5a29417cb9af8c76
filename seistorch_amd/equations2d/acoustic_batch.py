"""Uneven multi-model batches of the acoustic PML loop (data generation).

The reference runs every model through its own time loop (seistorch/rnn.py:
178-212) and boundary-saving backward (seistorch/checkpoint_new.py:146-217).
Here many independent problems, each with its own grid, shots, time axis,
dt, h, source and receivers, are advanced together: every time step is one
kernel launch over the tiles of all instances still running, and the whole
loop can be recorded into one hipGraph.  Per instance the records and the
gradient are bit-identical to ``acoustic.propagate`` with the same launch
shape and snapshot spacing.

    insts = [BatchInstance(vp, d, wavelet, src_zx, rec_flat, B, dt, h), ...]
    records = propagate_batch(insts)            # list of (nt_q, R_q)
    loss = sum(l2(r, obs) for r, obs in zip(records, observed))
    loss.backward()                             # vp.grad of every instance
"""
from dataclasses import dataclass

import torch

from .. import _lib
from .acoustic import NPML, SNAP_HBM_FRACTION, SNAP_MIN_EVERY, _f, ring_len

SHAPE_COLS = 5  # nz, nx, B, nt, R


@dataclass
class BatchInstance:
    """One problem: vp, d (nz, nx); wavelet (nt,); src_zx (B, 2) int32 padded
    (z, x); rec_flat (R,) int64 flat indices into the instance's (B, nz, nx)."""
    vp: torch.Tensor
    d: torch.Tensor
    wavelet: torch.Tensor
    src_zx: torch.Tensor
    rec_flat: torch.Tensor
    B: int
    dt: float
    h: float


def _snapshot_count(nt, K):
    return int(_lib.load().seis_snapshot_count(nt, K))


class Layout:
    """Offsets of every instance in the concatenated arenas (caller's order),
    as the library's batch entry points expect them."""

    def __init__(self, sizes, S, K, need_ring):
        self.sizes = sizes  # list of (nz, nx, B, nt, R)
        names = ("model", "wav", "src", "rec", "out", "ring", "snap", "field", "gpart")
        self.off = {k: [] for k in names}
        tot = dict.fromkeys(names, 0)
        for nz, nx, B, nt, R in sizes:
            plane = nz * nx
            G = -(-B // S)
            L = ring_len(nz, nx) if need_ring else 0
            size = {"model": plane, "wav": nt, "src": B, "rec": R, "out": nt * R,
                    "ring": (nt - 1) * B * L if (need_ring and nt > 1) else 0,
                    "snap": _snapshot_count(nt, K) * 2 * B * plane if K > 0 else 0,
                    "field": B * plane, "gpart": G * plane}
            for k in names:
                self.off[k].append(tot[k])
                tot[k] += size[k]
        self.total = tot


def default_vec_x(sizes):
    """Widest cells-per-lane that divides every instance's nx."""
    for v in (4, 2):
        if all(nx % v == 0 for _, nx, *_ in sizes):
            return v
    return 1


def default_snapshot_every(sizes, device):
    """One spacing for the whole batch: the densest (>= SNAP_MIN_EVERY) whose
    snapshots fit in SNAP_HBM_FRACTION of device memory."""
    total = torch.cuda.get_device_properties(device).total_memory
    need = sum(2 * B * nz * nx * 4 * max(nt - 2, 0) for nz, nx, B, nt, _ in sizes)
    if need == 0 or max(nt for *_, nt, _ in sizes) <= SNAP_MIN_EVERY + 1:
        return 0
    return max(SNAP_MIN_EVERY, -(-need // int(SNAP_HBM_FRACTION * total)))


def _check(insts, need_grad):
    """Host-side bounds checks for every instance (the kernels index unchecked)."""
    bad = []
    for q, it in enumerate(insts):
        if it.vp.dim() != 2 or it.d.shape != it.vp.shape:
            raise RuntimeError(f"instance {q}: vp {tuple(it.vp.shape)} and d "
                               f"{tuple(it.d.shape)} must be equal (nz, nx)")
        if it.wavelet.dim() != 1 or it.wavelet.shape[0] < 1 or it.wavelet.dtype != torch.float32:
            raise RuntimeError(f"instance {q}: wavelet must be a non-empty float32 vector")
        if it.src_zx.dtype != torch.int32 or tuple(it.src_zx.shape) != (it.B, 2):
            raise RuntimeError(f"instance {q}: src_zx must be int32 ({it.B}, 2)")
        if it.rec_flat.dtype != torch.int64 or it.rec_flat.dim() != 1:
            raise RuntimeError(f"instance {q}: rec_flat must be an int64 vector")
        nz, nx = it.vp.shape
        flags = [it.src_zx.min() < 0, it.src_zx[:, 0].max() >= nz, it.src_zx[:, 1].max() >= nx]
        if it.rec_flat.numel():
            flags += [it.rec_flat.min() < 0, it.rec_flat.max() >= it.B * nz * nx]
        if need_grad:
            m = NPML + 1
            if nz <= 2 * m or nx <= 2 * m:
                raise RuntimeError(f"instance {q}: grid {nz}x{nx} has no interior inside the "
                                   "50-cell boundary")
            flags.append(it.d[m:nz - m, m:nx - m].abs().max() != 0)
        bad.append(torch.stack([torch.as_tensor(f, device=it.vp.device) for f in flags]).any())
    bad = torch.stack(bad).cpu()
    if bool(bad.any()):
        q = int(bad.nonzero()[0, 0])
        raise RuntimeError(f"instance {q}: a source or receiver lies outside its grid, or d is "
                           "nonzero inside the 50-cell boundary")


def _host_tables(sizes, dts):
    shape = torch.tensor(sizes, dtype=torch.int64).reshape(-1)
    dt_h = torch.tensor(dts, dtype=torch.float32).reshape(-1)
    return shape, dt_h


def _hptr(t):
    return t.data_ptr()


class _BatchLoop(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp_arena, d_arena, wav_arena, src_arena, rec_arena, sizes, dts, launch,
                snapshot_every, use_graph, two_step=True):
        S, rz, vx = launch
        dev = vp_arena.device
        need_grad = ctx.needs_input_grad[0]
        K = 0
        if need_grad:
            K = default_snapshot_every(sizes, dev) if snapshot_every is None \
                else int(snapshot_every)
        lay = Layout(sizes, S, K, need_grad)
        shape, dt_h = _host_tables(sizes, dts)
        n = len(sizes)
        f32 = dict(device=dev, dtype=torch.float32)
        rec = torch.empty(lay.total["out"], **f32)
        u_a = torch.empty(lay.total["field"], **f32)
        u_b = torch.empty_like(u_a)
        # two steps per launch need the third and fourth field arenas
        u_c, u_d = (torch.empty_like(u_a), torch.empty_like(u_a)) if two_step else (None, None)
        ring = torch.empty(lay.total["ring"], **f32) if need_grad and lay.total["ring"] else None
        snaps = torch.empty(lay.total["snap"], **f32) if K > 0 and lay.total["snap"] else None
        table = torch.empty(n * _lib.BATCH_COLS, device=dev, dtype=torch.int64)
        _lib.call("seis_acoustic_forward_batch", n, _hptr(shape), _hptr(dt_h), _lib.ptr(vp_arena),
                  _lib.ptr(d_arena), _lib.ptr(wav_arena), _lib.ptr(src_arena),
                  _lib.ptr(rec_arena), _lib.ptr(rec), _lib.ptr(ring), _lib.ptr(snaps),
                  K if snaps is not None else 0, _lib.ptr(u_a), _lib.ptr(u_b), _lib.ptr(u_c),
                  _lib.ptr(u_d), _lib.ptr(table), NPML, S, rz, vx, int(use_graph),
                  _lib.stream_of(vp_arena))
        del u_c, u_d
        if need_grad:
            ctx.save_for_backward(vp_arena, d_arena, wav_arena, src_arena, rec_arena)
            ctx.state = (ring, snaps, u_a, u_b, table)
            ctx.meta = (sizes, dts, launch, K if snaps is not None else 0, use_graph, lay)
        # one output per instance, so the backward receives each cotangent once
        return tuple(rec[o:o + nt * R].view(nt, R)
                     for o, (_, _, _, nt, R) in zip(lay.off["out"], sizes))

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, *grad_recs):
        grad_rec = torch.cat([g.reshape(-1) for g in grad_recs])
        vp_arena, d_arena, wav_arena, src_arena, rec_arena = ctx.saved_tensors
        ring, snaps, u_a, u_b, table = ctx.state
        ctx.state = None  # the backward consumes the final forward state
        sizes, dts, (S, rz, vx), K, use_graph, lay = ctx.meta
        shape, dt_h = _host_tables(sizes, dts)
        mu_a = torch.empty_like(u_a)
        mu_b = torch.empty_like(u_a)
        gpart = torch.empty(lay.total["gpart"], device=u_a.device, dtype=torch.float32)
        grad = torch.empty_like(vp_arena)
        _lib.call("seis_acoustic_backward_batch", len(sizes), _hptr(shape), _hptr(dt_h),
                  _lib.ptr(vp_arena), _lib.ptr(d_arena), _lib.ptr(wav_arena),
                  _lib.ptr(src_arena), _lib.ptr(rec_arena),
                  _lib.ptr(grad_rec.contiguous().float()), _lib.ptr(ring), _lib.ptr(snaps), K,
                  _lib.ptr(u_a), _lib.ptr(u_b), _lib.ptr(mu_a), _lib.ptr(mu_b), _lib.ptr(gpart),
                  _lib.ptr(grad), _lib.ptr(table), NPML, S, rz, vx, int(use_graph),
                  _lib.stream_of(vp_arena))
        return (grad,) + (None,) * 10


def _launch_params(sizes, shape):
    from .. import autotune
    tuned = autotune.lookup_batch(sizes)
    default = tuned if tuned else (4, 2, default_vec_x(sizes))
    shape = tuple(shape) if shape is not None else (None, None, None)
    return tuple(int(v if v is not None else dv) for v, dv in zip(shape, default))


def propagate_batch(insts, shots_per_group=None, rows_per_thread=None, vec_x=None,
                    snapshot_every=None, use_graph=True, two_step=True):
    """Records of every instance, [(nt_q, R_q)], differentiable w.r.t. each vp.

    shots_per_group / rows_per_thread / vec_x: launch shape shared by the whole
    batch (None: tuned or default; vec_x must divide every nx).  snapshot_every:
    one spacing for all instances (None automatic, 0 ring only).  use_graph:
    record the time loop into one hipGraph (False: direct launches).
    two_step: the forward advances two time steps per launch (two more field
    arenas; records, rings and snapshots bit-identical to one step per launch).
    """
    insts = list(insts)
    if not insts:
        return []
    vps = [it.vp for it in insts]
    _lib.require_device(*vps)
    need_grad = torch.is_grad_enabled() and any(v.requires_grad for v in vps)
    _check(insts, need_grad)
    sizes = [(int(it.vp.shape[0]), int(it.vp.shape[1]), int(it.B), int(it.wavelet.shape[0]),
              int(it.rec_flat.shape[0])) for it in insts]
    dts = [(_f(it.dt), _f(it.h)) for it in insts]
    launch = _launch_params(sizes, (shots_per_group, rows_per_thread, vec_x))
    vp_arena = torch.cat([v.reshape(-1) for v in vps])
    with torch.no_grad():
        d_arena = torch.cat([it.d.reshape(-1).float() for it in insts])
        wav_arena = torch.cat([it.wavelet.reshape(-1) for it in insts])
        src_arena = torch.cat([it.src_zx.reshape(-1) for it in insts])
        rec_arena = torch.cat([it.rec_flat.reshape(-1) for it in insts])
    return list(_BatchLoop.apply(vp_arena, d_arena, wav_arena, src_arena, rec_arena, sizes,
                                 dts, launch, snapshot_every, bool(use_graph), bool(two_step)))
