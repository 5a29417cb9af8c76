"""Launch-shape autotuner for the fused time loops.

For a given equation and problem shape (shots, nz, nx) it times the fused
forward + adjoint over a short time axis for every candidate launch shape
-- (shots_per_group, rows_per_thread, vec_x) for the acoustic PML loop and
the uneven batch, shots_per_group for the elastic and HABC loops -- and keeps
the fastest.  Results live in a
process-wide table (and optionally a JSON file), keyed by shape, and
``propagate`` picks them up when the caller does not fix the launch shape.
"""
import itertools
import json
import os

import torch

_TABLE = {}
CANDIDATES_S = (2, 4, 8)
CANDIDATES_RZ = (1, 2, 4)
CANDIDATES_VX = (1, 2, 4)


def _key(eq, B, nz, nx):
    return f"{eq}:{B}:{nz}:{nx}"


def lookup(eq, B, nz, nx):
    return _TABLE.get(_key(eq, B, nz, nx))


def record(eq, B, nz, nx, shape):
    _TABLE[_key(eq, B, nz, nx)] = tuple(int(v) for v in shape)


def save(path):
    with open(path, "w") as f:
        json.dump({k: list(v) for k, v in _TABLE.items()}, f, indent=1)


def load(path):
    if os.path.exists(path):
        with open(path) as f:
            for k, v in json.load(f).items():
                _TABLE[k] = tuple(v)


def tune_acoustic(vp, d, wavelet, src_zx, rec_flat, B, dt, h, nt_probe=24, repeats=2,
                  candidates_s=CANDIDATES_S, candidates_rz=CANDIDATES_RZ,
                  candidates_vx=CANDIDATES_VX, verbose=False):
    """Time forward+adjoint for each launch shape on the real grid; keep the best."""
    from .equations2d import acoustic as ac
    nz, nx = vp.shape
    w = wavelet[:nt_probe].contiguous()
    results = {}
    for S, rz, vx in itertools.product(candidates_s, candidates_rz, candidates_vx):
        if S > B or nx % vx:
            continue
        best = float("inf")
        for _ in range(repeats):
            v = vp.detach().clone().requires_grad_(True)
            torch.cuda.synchronize()
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record()
            rec = ac.propagate(v, d, w, src_zx, rec_flat, B, dt, h, shots_per_group=S,
                               snapshot_every=0, rows_per_thread=rz, vec_x=vx)
            rec.backward(torch.ones_like(rec))
            t1.record()
            torch.cuda.synchronize()
            best = min(best, t0.elapsed_time(t1))
        results[(S, rz, vx)] = best
        if verbose:
            print(f"[autotune] acoustic S={S} rz={rz} vx={vx}: {best:.2f} ms", flush=True)
    shape = min(results, key=results.get)
    record("acoustic", B, nz, nx, shape)
    return shape, results


def _time_grad(run, repeats):
    """Best wall time (ms, HIP events) of run() = forward + backward."""
    best = float("inf")
    for _ in range(repeats):
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record()
        run()
        t1.record()
        torch.cuda.synchronize()
        best = min(best, t0.elapsed_time(t1))
    return best


def tune_elastic(vp, vs, rho, d, wavelet, src_zx, rec_flat, B, dt, h, nt_probe=16, repeats=2,
                 candidates_s=(1, 2, 4, 8), verbose=False, **kw):
    """Shots per thread of the fused elastic forward and, separately, of its
    adjoint (they need not be equal): the forward alone, then forward +
    adjoint with the best forward, timed over nt_probe steps on the real grid
    for each candidate.  The pair (forward, adjoint) is recorded for
    ``elastic.propagate``."""
    from .equations2d import elastic as el
    nz, nx = vp.shape
    w = wavelet[:nt_probe].contiguous()
    cands = [S for S in candidates_s if S <= B]
    results = {}

    def fwd(S):
        with torch.no_grad():
            el.propagate(vp, vs, rho, d, w, src_zx, rec_flat, B, dt, h, shots_per_group=S, **kw)
    for S in cands:
        results[("fwd", S)] = _time_grad(lambda: fwd(S), repeats)
    sf = min(cands, key=lambda S: results[("fwd", S)])

    def grad(Sb):
        m = [x.detach().clone().requires_grad_(True) for x in (vp, vs, rho)]
        r = el.propagate(*m, d, w, src_zx, rec_flat, B, dt, h, shots_per_group=(sf, Sb),
                         snapshot_every=0, **kw)
        r.backward(torch.ones_like(r))
    for Sb in cands:
        results[("fwd+adj", Sb)] = _time_grad(lambda: grad(Sb), repeats)
    sb = min(cands, key=lambda S: results[("fwd+adj", S)])
    if verbose:
        for k, v in results.items():
            print(f"[autotune] elastic {k}: {v:.2f} ms", flush=True)
    record("elastic", B, nz, nx, (sf, sb))
    return (sf, sb), results


def tune_habc(vp, b, wavelet, src_zx, rec_flat, B, dt, h, nt_probe=16, repeats=2,
              candidates_s=(1, 2, 4, 8), verbose=False):
    """Shots per thread group of the fused HABC forward and adjoint, as
    tune_elastic; recorded for ``acoustic_habc.propagate``."""
    from .equations2d import acoustic_habc as ah
    nz, nx = vp.shape
    w = wavelet[:nt_probe].contiguous()
    results = {}
    for S in candidates_s:
        if S > B:
            continue

        def run():
            v = vp.detach().clone().requires_grad_(True)
            r = ah.propagate(v, b, w, src_zx, rec_flat, B, dt, h, shots_per_group=S,
                             snapshot_every=0)
            r.backward(torch.ones_like(r))
        results[(S,)] = _time_grad(run, repeats)
        if verbose:
            print(f"[autotune] habc S={S}: {results[(S,)]:.2f} ms", flush=True)
    shape = min(results, key=results.get)
    record("habc", B, nz, nx, shape)
    return shape, results


def _batch_key(sizes):
    import hashlib
    blob = json.dumps([list(map(int, s)) for s in sizes]).encode()
    return "acoustic_batch:" + hashlib.sha1(blob).hexdigest()[:16]


def lookup_batch(sizes):
    return _TABLE.get(_batch_key(sizes))


def tune_batch(insts, nt_probe=16, repeats=2, candidates_s=CANDIDATES_S,
               candidates_rz=CANDIDATES_RZ, candidates_vx=CANDIDATES_VX, verbose=False):
    """Time one batched forward+adjoint of every instance over nt_probe steps
    (all instances active, the dominant regime) for each launch shape."""
    from .equations2d import acoustic_batch as ab
    sizes = [(int(it.vp.shape[0]), int(it.vp.shape[1]), int(it.B), int(it.wavelet.shape[0]),
              int(it.rec_flat.shape[0])) for it in insts]
    probe = [ab.BatchInstance(it.vp.detach(), it.d, it.wavelet[:nt_probe].contiguous(),
                              it.src_zx, it.rec_flat, it.B, it.dt, it.h) for it in insts]
    bmin = min(s[2] for s in sizes)
    results = {}
    for S, rz, vx in itertools.product(candidates_s, candidates_rz, candidates_vx):
        if S > bmin or any(s[1] % vx for s in sizes):
            continue
        best = float("inf")
        for _ in range(repeats):
            for it in probe:
                it.vp = it.vp.detach().requires_grad_(True)
            torch.cuda.synchronize()
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record()
            recs = ab.propagate_batch(probe, shots_per_group=S, rows_per_thread=rz, vec_x=vx,
                                      snapshot_every=0, use_graph=False)
            torch.autograd.backward(recs, [torch.ones_like(r) for r in recs])
            t1.record()
            torch.cuda.synchronize()
            best = min(best, t0.elapsed_time(t1))
        results[(S, rz, vx)] = best
        if verbose:
            print(f"[autotune] batch S={S} rz={rz} vx={vx}: {best:.2f} ms", flush=True)
    shape = min(results, key=results.get)
    _TABLE[_batch_key(sizes)] = shape
    return shape, results
