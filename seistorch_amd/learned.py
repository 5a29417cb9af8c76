"""Training data and training loop of the learned preconditioner.

Pairs come from the uneven-batch generator (seistorch_amd.generator): for
every seeded model, the L2-misfit gradient at the smoothed start model
(computed by the batched solver) and the perturbation true - start.  A pair
is the whole unpadded model cut to multiples of 16 cells from its surface
corner, so that the network sees what it sees inside FWI (the surface
acquisition footprint and a per-image gradient normalisation): input =
(gradient with the top `mute` rows zeroed, start model), target =
(true - start) / vscale.  Mini-batches pad their members to a common size
(gradient and target zero, model edge-replicated) and mask the padding out of
the loss.
"""
import time

import numpy as np
import torch
import torch.nn.functional as F

from . import generator, precond

PAD = generator.PAD


def make_pairs(n_models, device, seed=20230503, mute=5, chunk=128, vscale=500.0, spec=None,
               log=None, first=0):
    """[(grad, model, target)] per model, each (h, w) FP32 on the device, for
    the generator's instances first .. first + n_models - 1 of the seed."""
    spec = spec or generator.GenSpec()
    spec = generator.GenSpec(**{**spec.__dict__, "n_models": n_models})
    host = generator.make_instances(seed=seed, spec=spec, first=first)
    out = []
    t0 = time.perf_counter()
    for c0 in range(0, len(host), chunk):
        part = generator.to_device(host[c0:c0 + chunk], device)
        obs = generator.observed(part)
        _, grads = generator.misfit_gradients(part, obs, snapshot_every=500)
        for it, g in zip(part, grads):
            h, w = it["nz_i"] // 16 * 16, it["nx_i"] // 16 * 16
            sl = (slice(PAD, PAD + h), slice(PAD, PAD + w))
            gi = g[sl].clone()
            gi[:mute] = 0.0
            vi = it["vp_start"][sl].clone()
            ti = (it["vp_true"][sl] - vi) / vscale
            out.append((gi, vi, ti))
        del part, obs, grads
        # the batch's ring and field buffers differ in size from chunk to chunk:
        # hand the cached blocks back so that the next chunk's large ring
        # buffer is not refused for fragmentation
        torch.cuda.empty_cache()
        if log:
            log(f"[pairs] {min(c0 + chunk, len(host))}/{len(host)} models, "
                f"{time.perf_counter() - t0:.1f} s")
    return out


def collate(pairs):
    """Pad a list of pairs to their largest (h, w): (grad, model, target, mask)."""
    H = max(p[0].shape[0] for p in pairs)
    W = max(p[0].shape[1] for p in pairs)
    n = len(pairs)
    dev = pairs[0][0].device
    G = torch.zeros((n, H, W), device=dev)
    T = torch.zeros((n, H, W), device=dev)
    Mk = torch.zeros((n, H, W), device=dev)
    Mo = torch.empty((n, H, W), device=dev)
    for i, (g, m, t) in enumerate(pairs):
        h, w = g.shape
        G[i, :h, :w] = g
        T[i, :h, :w] = t
        Mk[i, :h, :w] = 1.0
        Mo[i] = F.pad(m[None, None], (0, W - w, 0, H - h), mode="replicate")[0, 0]
    return G, Mo, T, Mk


def _flip(pair):
    return tuple(torch.flip(a, [1]) for a in pair)


def train(net, opt, pairs, steps, batch=8, seed=0, log=None, log_every=100, flip=False):
    """AdamW steps on random mini-batches of whole-model pairs; returns the
    (step, loss) history.  flip: each drawn pair is mirrored in x with
    probability 1/2 (the generator's acquisition, evenly spaced surface shots
    and a receiver at every surface cell, is symmetric under the mirror up to
    the rounding of the shot positions)."""
    gen = torch.Generator().manual_seed(seed)
    n = len(pairs)
    losses = []
    for s in range(steps):
        idx = torch.randint(0, n, (min(batch, n),), generator=gen).tolist()
        flips = torch.rand(len(idx), generator=gen).lt(0.5).tolist() if flip else [False] * len(idx)
        g, m, t, mk = collate([_flip(pairs[i]) if f else pairs[i] for i, f in zip(idx, flips)])
        sc = precond.input_scales(g, m)
        loss = precond.train_step(net, opt, g, m, sc, t, mk)
        if (s + 1) % log_every == 0 or s == steps - 1:
            losses.append((s + 1, float(loss)))
            if log:
                log(f"[train] step {s + 1}/{steps}: loss {losses[-1][1]:.4f}")
    return losses


def evaluate(net, pairs):
    """Mean squared error of the network and of the zero direction over the
    pairs (each model on its own, no padding)."""
    tot = zero = cells = 0.0
    for g, m, t in pairs:
        out = net.forward(g[None].contiguous(), m[None].contiguous(),
                          precond.input_scales(g[None], m[None]))[0]
        tot += float(((out - t) ** 2).sum())
        zero += float((t ** 2).sum())
        cells += t.numel()
    return tot / cells, zero / cells


def heldout_errors(net, pairs):
    """Squared errors, summed over held-out pairs, against target = true -
    start: (zero direction, best multiple of the max-normalised gradient,
    network, best multiple of the network's output).  The second is the
    plain method's update direction with its step chosen in hindsight, the
    fairest form of the method without the network."""
    z = gs = nn = nb = 0.0
    for g, m, t in pairs:
        o = net.forward(g[None].contiguous(), m[None].contiguous(),
                        precond.input_scales(g[None], m[None]))[0]
        gn = g / g.abs().max().clamp_min(1e-30)
        a = float((gn * t).sum() / (gn * gn).sum().clamp_min(1e-30))
        b = float((o * t).sum() / (o * o).sum().clamp_min(1e-30))
        z += float((t * t).sum())
        gs += float(((a * gn - t) ** 2).sum())
        nn += float(((o - t) ** 2).sum())
        nb += float(((b * o - t) ** 2).sum())
    return z, gs, nn, nb

