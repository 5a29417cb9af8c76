"""Training data and training loop of the learned preconditioner.

Pairs come from the uneven-batch generator (seistorch_amd.generator): for
every seeded model, the L2-misfit gradient at the smoothed start model
(computed by the batched solver) and the perturbation true - start.  Each
pair is cut into fixed-size crops of the unpadded grid (network input:
gradient with the top `mute` rows zeroed and the start model; target:
(true - start) / vscale), so that crops of models of different sizes batch
together.
"""
import time

import numpy as np
import torch

from . import generator, precond

PAD = generator.PAD


def make_pairs(n_models, device, seed=20230503, crop=128, crops_per_model=4, mute=5,
               chunk=256, vscale=500.0, spec=None, log=None):
    """(grad, model, target) crops as (M, crop, crop) FP32 device tensors."""
    spec = spec or generator.GenSpec()
    spec = generator.GenSpec(**{**spec.__dict__, "n_models": n_models})
    host = generator.make_instances(seed=seed, spec=spec)
    rng = np.random.default_rng(seed + 1)
    G, Mo, T = [], [], []
    t0 = time.perf_counter()
    for c0 in range(0, len(host), chunk):
        part = generator.to_device(host[c0:c0 + chunk], device)
        obs = generator.observed(part)
        _, grads = generator.misfit_gradients(part, obs, snapshot_every=500)
        for it, g in zip(part, grads):
            nz, nx = it["nz_i"], it["nx_i"]
            if nz < crop or nx < crop:
                continue
            gi = g[PAD:PAD + nz, PAD:PAD + nx].clone()
            gi[:mute] = 0.0
            vi = it["vp_start"][PAD:PAD + nz, PAD:PAD + nx]
            ti = (it["vp_true"][PAD:PAD + nz, PAD:PAD + nx] - vi) / vscale
            for _ in range(crops_per_model):
                z0 = int(rng.integers(0, nz - crop + 1))
                x0 = int(rng.integers(0, nx - crop + 1))
                sl = (slice(z0, z0 + crop), slice(x0, x0 + crop))
                G.append(gi[sl]); Mo.append(vi[sl]); T.append(ti[sl])
        del part, obs, grads
        if log:
            log(f"[pairs] {min(c0 + chunk, len(host))}/{len(host)} models, "
                f"{len(G)} crops, {time.perf_counter() - t0:.1f} s")
    return (torch.stack(G).contiguous(), torch.stack(Mo).contiguous(),
            torch.stack(T).contiguous())


def train(net, opt, pairs, steps, batch=16, seed=0, log=None, log_every=100):
    """AdamW steps on random mini-batches of the pairs; returns the loss history."""
    G, Mo, T = pairs
    n = G.shape[0]
    gen = torch.Generator().manual_seed(seed)
    losses = []
    for s in range(steps):
        idx = torch.randint(0, n, (min(batch, n),), generator=gen).to(G.device)
        g, m, t = G[idx].contiguous(), Mo[idx].contiguous(), T[idx].contiguous()
        sc = precond.input_scales(g, m)
        loss = precond.train_step(net, opt, g, m, sc, t)
        if (s + 1) % log_every == 0 or s == steps - 1:
            losses.append((s + 1, float(loss)))
            if log:
                log(f"[train] step {s + 1}/{steps}: loss {losses[-1][1]:.4f}")
    return losses


def evaluate(net, pairs, batch=32):
    """Mean loss of the network and of the zero direction over the pairs."""
    G, Mo, T = pairs
    tot = zero = 0.0
    for i in range(0, G.shape[0], batch):
        g, m, t = G[i:i + batch], Mo[i:i + batch], T[i:i + batch]
        out = net.forward(g, m, precond.input_scales(g, m))
        tot += float(((out - t) ** 2).sum())
        zero += float((t ** 2).sum())
    return tot / G.numel(), zero / G.numel()
