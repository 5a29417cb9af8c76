"""Single-GPU acoustic FWI driver: the reference's multiscale loop
(``run_multiscale``: band-filtered records, Adam or max-normalised steepest
descent per frequency scale), and plain-gradient iterations against
iterations preconditioned by the learned U-Net, compared at equal wall-clock
time (``run_plain`` / ``run_preconditioned``).

The plain method follows the reference's FWI loop (seistorch/fwi.py and
seistorch_dist.py:166-303: forward, L2 misfit of loss.py:409-421, backward,
model update) with a steepest-descent step normalised to a velocity change
and accepted only if the misfit decreases (otherwise halved).  The gradient
is muted within `mute` cells of the surface, as the reference's seabed cut
(seistorch/process.py:43-47) does, and only the unpadded model is updated;
the boundary strips are re-padded from its edges (geom.py:213-240).

The preconditioned method feeds (gradient, model) to the network, which
predicts the model perturbation true - start; in its first iteration(s) the
trial update is that perturbation plus the plain step, accepted if the
misfit decreases, otherwise the plain step of the same iteration is taken
(the fallback); later iterations are plain.  Applying the network at every
iteration was measured to drive the model away (it was trained on gradients
at smoothed start models only).
"""
import time
from dataclasses import dataclass, field

import numpy as np
import torch
import torch.nn.functional as F

from .equations2d import acoustic as ac

PAD = 50


@dataclass
class Problem:
    vp_true: torch.Tensor   # padded (nz, nx)
    vp_init: torch.Tensor   # padded (nz, nx)
    d: torch.Tensor
    wavelet: torch.Tensor
    src: torch.Tensor
    rec: torch.Tensor
    B: int
    dt: float
    h: float
    obs: torch.Tensor = None
    mute: int = 5
    vmin: float = 1400.0
    vmax: float = 5000.0
    band: object = "all"     # multiscale frequency spec (seistorch/signal.py:25-37)
    filter_ord: int = 3      # Butterworth order (template.yml:44)
    _fobs: tuple = None      # (band, filtered observed records)
    # the configuration's `equation:` (model.py:64-77): "acoustic" (PML, d is
    # the damping) or "acoustic_habc" (d is the HABC weight field)
    equation: str = "acoustic"

    def interior(self, v):
        return v[PAD:v.shape[0] - PAD, PAD:v.shape[1] - PAD]

    def pad(self, vi):
        return F.pad(vi[None, None], (PAD, PAD, PAD, PAD), mode="replicate")[0, 0].contiguous()

    def _propagate(self, vp, illum=None):
        if self.equation == "acoustic":
            return ac.propagate(vp, self.d, self.wavelet, self.src, self.rec, self.B, self.dt,
                                self.h, illum=illum)
        if self.equation == "acoustic_habc":
            from .equations2d import acoustic_habc as ah
            return ah.propagate(vp, self.d, self.wavelet, self.src, self.rec, self.B, self.dt,
                                self.h, illum=illum)
        raise ValueError(f"unsupported equation {self.equation!r}")

    def records(self, vp):
        with torch.no_grad():
            return self._propagate(vp)

    def _filter(self, rec):
        """Records filtered to the current band (seistorch_dist.py:222-225:
        synthetic and observed records through the same zero-phase filter);
        differentiable."""
        if self.band == "all":
            return rec
        from . import signal as sg
        b, a = sg.butter(self.filter_ord, self.band, self.dt)
        return sg.filtfilt(rec, b, a)

    def observed(self):
        if self.band == "all":
            return self.obs
        if self._fobs is None or self._fobs[0] != self.band:
            with torch.no_grad():
                self._fobs = (self.band, self._filter(self.obs))
        return self._fobs[1]

    def misfit(self, vp):
        with torch.no_grad():
            return float(torch.sum((self._filter(self.records(vp)) - self.observed()) ** 2))

    def gradient(self, vp, illum=None, padded=False):
        """(misfit, gradient): the interior gradient, or with padded=True the
        gradient over the padded grid (zero in the boundary strips, which the
        boundary-saving adjoint does not rebuild); the top `mute` interior
        rows are zeroed.  illum: optional (nz, nx) buffer that receives the
        forward's source illumination."""
        v = vp.detach().clone().requires_grad_(True)
        rec = self._propagate(v, illum=illum)
        loss = torch.sum((self._filter(rec) - self.observed()) ** 2)
        loss.backward()
        g = v.grad.clone()
        if self.mute:
            g[PAD:PAD + self.mute] = 0.0
        return float(loss), (g if padded else self.interior(g).clone())

    def rmse(self, vp):
        return float(torch.sqrt(torch.mean((self.interior(vp) - self.interior(self.vp_true)) ** 2)))


@dataclass
class History:
    t: list = field(default_factory=list)
    misfit: list = field(default_factory=list)
    rmse: list = field(default_factory=list)
    kind: list = field(default_factory=list)

    def add(self, t, misfit, rmse, kind):
        self.t.append(t); self.misfit.append(misfit); self.rmse.append(rmse); self.kind.append(kind)


def _clip(p, vi):
    return vi.clamp(p.vmin, p.vmax)


def plain_step(p, vp, loss, g, step):
    """Steepest descent: dv = -step * g / max|g|; halve until the misfit drops
    (at most 3 tries).  Returns (vp, misfit, step, accepted)."""
    gmax = float(g.abs().max())
    if gmax == 0.0:
        return vp, loss, step, False
    vi = p.interior(vp)
    for _ in range(3):
        trial = p.pad(_clip(p, vi - step * g / gmax))
        m = p.misfit(trial)
        if m < loss:
            return trial, m, step, True
        step *= 0.5
    return vp, loss, step, False


def descent_direction(p, vp, smooth=None, source_illumination=False):
    """(misfit, raw interior gradient, update direction over the interior).
    Without smooth and source_illumination the direction is the raw gradient
    (top `mute` rows zeroed); with them it is the reference's post-processing
    of the padded gradient before its optimizer step (seistorch_dist.py:270-280
    through postprocess_padded: `counts` Gaussian passes, then the division by
    the forward's source illumination, rnn.py:127-128,204-205), muted again."""
    if smooth is None and not source_illumination:
        loss, g = p.gradient(vp)
        return loss, g, g
    illum = torch.empty_like(p.vp_true) if source_illumination else None
    loss, gp = p.gradient(vp, illum=illum, padded=True)
    raw = p.interior(gp).clone()
    d = postprocess_padded(gp, smooth, None, illum)
    if p.mute:
        d[:p.mute] = 0.0
    return loss, raw, d


def run_plain(p, iters, step0=40.0, time_budget=None, log=None, smooth=None,
              source_illumination=False):
    """Steepest descent along descent_direction (plain_step's line search)."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    vp = p.vp_init.clone()
    hist = History()
    hist.add(0.0, p.misfit(vp), p.rmse(vp), "start")
    step = step0
    for it in range(iters):
        loss, _, g = descent_direction(p, vp, smooth, source_illumination)
        vp, m, step, ok = plain_step(p, vp, loss, g, step)
        if ok:
            step *= 1.25
        torch.cuda.synchronize()
        hist.add(time.perf_counter() - t0, m, p.rmse(vp), "plain" if ok else "none")
        if log:
            log(f"[plain] it {it + 1}: misfit {m:.4e} rmse {hist.rmse[-1]:.1f} step {step:.1f}")
        if time_budget and hist.t[-1] >= time_budget:
            break
    return vp, hist


def run_preconditioned(p, net, iters, vscale=500.0, step0=40.0, net_iters=1, time_budget=None,
                       log=None, smooth=None, source_illumination=False, scales=(1.0, 0.5)):
    """FWI whose first `net_iters` iterations try the network's predicted
    perturbation (from the raw gradient and the model, the inputs it was
    trained on) plus the plain step, at each of `scales` of the prediction in
    turn, accepting the first trial that lowers the misfit; otherwise (the
    fallback) the plain step of the same iteration along descent_direction.
    Later iterations are plain.

    The network was trained on gradients at smoothed start models, so it is
    consulted where its inputs look like its training data; afterwards the
    plain steepest-descent iterations continue from the improved model."""
    from . import precond
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    vp = p.vp_init.clone()
    hist = History()
    hist.add(0.0, p.misfit(vp), p.rmse(vp), "start")
    step = step0
    for it in range(iters):
        loss, raw, g = descent_direction(p, vp, smooth, source_illumination)
        kind, m = "none", loss
        if it < net_iters:
            vi = p.interior(vp).contiguous()
            sc = precond.input_scales(raw[None], vi[None])
            d = net.forward(raw[None].contiguous(), vi[None], sc)[0] * vscale
            d[:p.mute] = 0.0
            gmax = float(g.abs().max())
            plain_dv = -step * g / gmax if gmax > 0.0 else torch.zeros_like(g)
            for a in scales:
                trial = p.pad(_clip(p, vi + a * d + plain_dv))
                mt = p.misfit(trial)
                if mt < loss:
                    vp, m, kind = trial, mt, f"net*{a:g}+plain"
                    step *= 1.25
                    break
        if kind == "none":  # the plain step of this iteration (the fallback)
            vp, m, step, ok = plain_step(p, vp, loss, g, step)
            kind = "plain" if ok else "none"
            if ok:
                step *= 1.25
        torch.cuda.synchronize()
        hist.add(time.perf_counter() - t0, m, p.rmse(vp), kind)
        if log:
            log(f"[precond] it {it + 1}: {kind} misfit {m:.4e} rmse {hist.rmse[-1]:.1f}")
        if time_budget and hist.t[-1] >= time_budget:
            break
    return vp, hist


def misfit_at(hist, t):
    """Best misfit reached by wall-clock time t."""
    best = None
    for ti, m in zip(hist.t, hist.misfit):
        if ti <= t:
            best = m if best is None else min(best, m)
    return best


def time_to(hist, target):
    """Wall-clock time at which the misfit first reaches `target` (None if never)."""
    for ti, m in zip(hist.t, hist.misfit):
        if m <= target:
            return ti
    return None


def _gd_direction(g, grad_clamp=True):
    """seistorch/optimizer.py:124-142 (Steepestdescent): the gradient clamped
    to its 2 % and 98 % quantiles, scaled so that its largest entry is 1 (the
    update is then lr times this)."""
    if grad_clamp:
        q = torch.quantile(g.flatten().float(),
                           torch.tensor([0.02, 0.98], device=g.device, dtype=torch.float32))
        g = torch.clamp(g, min=float(q[0]), max=float(q[1]))
    m = float(g.abs().max())
    return g / m if m > 0 else g


def postprocess_padded(g_pad, smooth=None, seabed=None, illum=None, multiple=False):
    """The reference's post-processing of a gradient over the padded grid
    (seistorch_dist.py:270-280 with process.py:26-117): the seabed mask of
    the unpadded grid is zero-padded (no top pad with a free surface,
    process.py:26-40); then g * mask, `counts` Gaussian passes over the padded
    gradient (smooth: cfg['training']['smooth']), g * mask, g / illum -- each
    step when its argument is given.  Returns the interior, the part the loop
    updates: cells near its edges see the strips' values through the
    smoothing, as in the reference, not a reflection at the interior edge."""
    from . import process
    mask = None
    top = 0 if multiple else PAD
    if seabed is not None:
        mask = F.pad(seabed[None, None].float(), (PAD, PAD, top, PAD), mode="constant",
                     value=0.0)[0, 0].contiguous()
    if smooth is not None or mask is not None or illum is not None:
        g_pad = process.smooth_gradient(g_pad.contiguous(), smooth,
                                        pre_mask=mask if smooth is not None else None,
                                        post_mask=mask, illum=illum)
    return g_pad[top:g_pad.shape[0] - PAD, PAD:g_pad.shape[1] - PAD].contiguous()


def run_multiscale(p, scales, epochs, lr=5.0, optimizer="adam", scale_decay=0.85,
                   lr_decay=0.995, grad_clamp=True, smooth=None, seabed=None,
                   source_illumination=False, log=None):
    """The reference's FWI loop on one GPU (seistorch_dist.py:166-303): for
    each frequency scale (seistorch/signal.py band spec, or "all") a fresh
    optimizer with lr * scale_decay**scale (setup.py:157-183), `epochs`
    iterations of forward, band-filtered L2 misfit, adjoint gradient (seabed
    rows muted) and update, the learning rate decayed by lr_decay per epoch
    (ExponentialLR).  optimizer "adam": torch.optim.Adam with eps = 1e-22
    (optimizer.py:8-15); "steepestdescent": v -= lr * g / max|g| with the
    2-98 % quantile clamp (optimizer.py:124-142).  The unpadded model is the
    optimised variable; the boundary strips are re-padded from its edges after
    every update (process.py:50-63).  smooth (cfg['training']['smooth']:
    counts, sigma and radius per axis) and seabed (a mask of the interior's
    shape) switch on --grad-smooth and --grad-cut (seistorch_dist.py:270-275),
    source_illumination --source-illumination (:279-280: the gradient divided
    by the forward's illumination), on the HIP post-processing kernels,
    applied to the padded gradient before the interior is updated
    (postprocess_padded).  Returns (vp, History)."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    vi = p.interior(p.vp_init).clone().contiguous()
    vp = p.pad(vi)
    hist = History()
    band0 = p.band
    try:
        for si, band in enumerate(scales):
            p.band = band
            slr = lr * scale_decay ** si
            x = torch.nn.Parameter(vi.clone())
            opt = torch.optim.Adam([x], lr=slr, eps=1e-22) if optimizer == "adam" else None
            sched = torch.optim.lr_scheduler.ExponentialLR(opt, lr_decay) if opt else None
            cur = slr
            for ep in range(epochs):
                illum = torch.empty_like(p.vp_true) if source_illumination else None
                loss, g = p.gradient(p.pad(x.detach()), illum=illum, padded=True)
                g = postprocess_padded(g, smooth, seabed, illum)
                if opt is not None:
                    opt.zero_grad()
                    x.grad = g.contiguous()
                    opt.step()
                    sched.step()
                else:
                    with torch.no_grad():
                        x -= cur * _gd_direction(g, grad_clamp)
                    cur *= lr_decay
                with torch.no_grad():
                    x.clamp_(p.vmin, p.vmax)
                torch.cuda.synchronize()
                hist.add(time.perf_counter() - t0, loss, p.rmse(p.pad(x.detach())),
                         f"scale{si}")
                if log:
                    log(f"[multiscale] band {band} epoch {ep + 1}: misfit {loss:.4e} "
                        f"rmse {hist.rmse[-1]:.1f}")
            vi = x.detach().contiguous()
            vp = p.pad(vi)
    finally:
        p.band = band0
    return vp, hist


# ------------------------------------------------ configuration-driven runs

def problem_from_config(cfg, device="cuda", shots=None, mute=0):
    """A Problem from the reference's configuration keys (template.yml;
    seistorch_dist.py:90-160): ``equation`` ("acoustic" with a PML boundary,
    "acoustic_habc", "elastic": an ElasticProblem with vp, vs and rho,
    geom.invlist, source_type and receiver_type, or "acoustic_rho_habc": a
    RhoProblem with vp and rho), geom.truePath / initPath ``vp`` (.npy (nz, nx)
    models of the unpadded grid), geom.sources / receivers (.pkl),
    geom.obsPath (.npy records, one (nt, nr, 1) array per shot), the Ricker
    wavelet from geom.fm / dt / nt / wavelet_delay / wavelet_inverse or a
    geom.wavelet .npy (setup.py:232-252), geom.h, geom.boundary.width and
    training.filter_ord.  The fused path's ring sits at index 49, so the
    boundary width must be 50; multiples (no top boundary) are not supported
    here.  shots: the shot numbers to use (all by default).  mute: interior
    rows below the surface whose gradient is zeroed (0: none, as the
    reference's loop, which cuts only through the seabed mask)."""
    from . import io as sio
    from .habc import generate_habc_coefficients_2d
    from .pml import generate_pml_coefficients_2d
    from .synthetic import rec_flat_index
    from .utils import pad_model, ricker_wave
    geom = cfg["geom"]
    width = int(geom["boundary"]["width"])
    if width != PAD:
        raise ValueError(f"boundary width {width}: the fused path needs {PAD}")
    if geom.get("multiple", False):
        raise NotImplementedError("multiples (free surface without top boundary) are not "
                                  "supported by problem_from_config")
    equation = cfg.get("equation", "acoustic")
    if equation not in ("acoustic", "acoustic_habc", "elastic", "acoustic_rho_habc"):
        raise ValueError(f"unsupported equation {equation!r}")
    dt, h, nt = float(geom["dt"]), float(geom["h"]), int(geom["nt"])

    def model(key, name="vp"):
        path = (geom.get(key) or {}).get(name)
        return None if not path else torch.from_numpy(
            pad_model(sio.load_model(path), width).astype("float32")).to(device)

    vp_init = model("initPath")
    vp_true = model("truePath")
    if vp_init is None:
        raise ValueError("geom.initPath.vp is required")
    if vp_true is None:
        vp_true = vp_init.clone()
    nz, nx = vp_init.shape
    if geom.get("wavelet"):
        w = torch.from_numpy(sio.load_npy(geom["wavelet"]).astype("float32"))
    else:
        w = ricker_wave(geom["fm"], dt, nt, geom.get("wavelet_delay", 80),
                        inverse=bool(geom.get("wavelet_inverse", False)))
    if w.shape[0] != nt:
        raise ValueError(f"wavelet has {w.shape[0]} samples, geom.nt is {nt}")
    src_list = sio.read_pkl(geom["sources"])
    rec_list = sio.read_pkl(geom["receivers"])
    shots = list(range(len(src_list))) if shots is None else list(shots)
    src, rec = sio.setup_acquisition(src_list, rec_list, width, False, shots)
    if equation in ("acoustic_habc", "acoustic_rho_habc"):
        d = generate_habc_coefficients_2d((nz, nx), width)
    else:
        d = generate_pml_coefficients_2d((nz, nx), width)
    common = (d.to(device), w.to(device), torch.from_numpy(src).to(device),
              torch.from_numpy(rec_flat_index(rec, nz, nx)).to(device), len(shots), dt, h)
    filter_ord = int(cfg.get("training", {}).get("filter_ord", 3))
    if equation == "elastic":
        true, init = {}, {}
        for k in ("vp", "vs", "rho"):
            init[k] = model("initPath", k)
            true[k] = model("truePath", k)
            if init[k] is None:
                raise ValueError(f"geom.initPath.{k} is required for the elastic equation")
            if true[k] is None:
                true[k] = init[k].clone()
        inv = geom.get("invlist") or {"vp": True, "vs": True}
        p = ElasticProblem(true, init, *common,
                           src_type=tuple(geom.get("source_type") or ("txx", "tzz")),
                           rec_type=tuple(geom.get("receiver_type") or ("vx", "vz")),
                           invlist=tuple(k for k in ("vp", "vs", "rho") if inv.get(k)),
                           filter_ord=filter_ord, mute=mute)
        nch = len(p.rec_type)
    elif equation == "acoustic_rho_habc":
        true, init = {}, {}
        for k in ("vp", "rho"):
            init[k] = model("initPath", k)
            true[k] = model("truePath", k)
            if init[k] is None:
                raise ValueError(f"geom.initPath.{k} is required for acoustic_rho_habc")
            if true[k] is None:
                true[k] = init[k].clone()
        inv = geom.get("invlist") or {"vp": True, "rho": True}
        p = RhoProblem(true, init, *common,
                       invlist=tuple(k for k in ("vp", "rho") if inv.get(k)),
                       filter_ord=filter_ord, mute=mute)
        nch = 1
    else:
        p = Problem(vp_true, vp_init, *common, filter_ord=filter_ord, equation=equation,
                    mute=mute)
        nch = 1
    if geom.get("obsPath"):
        records = sio.SeisRecord(cfg)
        records.setup("inversion")
        obs = sio.stack_records([records[i] for i in shots])
        want = (nt, len(p.rec), nch)
        if obs.shape != want:
            raise ValueError(f"observed records {obs.shape} do not match (nt, receivers, "
                             f"channels) {want}")
        p.obs = torch.from_numpy(obs if nch > 1 else obs[..., 0]).to(device)
    return p


def run_from_config(cfg, device="cuda", shots=None, optimizer="adam", grad_smooth=False,
                    grad_cut=False, source_illumination=False, log=None, mute=0):
    """The multiscale FWI loop of seistorch_dist.py:166-303 driven by a
    configuration: geom.multiscale bands, training.N_epochs, lr (a number or
    {"vp": lr}), scale_decay, lr_decay; the command-line switches
    --grad-smooth (training.smooth), --grad-cut (geom.seabed, an .npy mask of
    the unpadded grid) and --source-illumination.
    Returns (Problem, vp, History); for the elastic and variable-density
    equations (ElasticProblem / RhoProblem, {name: padded model}, History)."""
    from . import io as sio
    p = problem_from_config(cfg, device, shots, mute=mute)
    if p.obs is None:
        raise ValueError("geom.obsPath is required for an inversion")
    tr = cfg.get("training", {})
    lr = tr.get("lr", 5.0)
    scales = cfg["geom"].get("multiscale") or ["all"]
    seabed = None
    if grad_cut:
        seabed = torch.from_numpy(sio.load_model(cfg["geom"]["seabed"])).to(device)
    if isinstance(p, ElasticProblem):
        models, hist = run_multiscale_elastic(p, scales, int(tr["N_epochs"]), lr,
                                              optimizer=optimizer,
                                              scale_decay=float(tr.get("scale_decay", 0.85)),
                                              lr_decay=float(tr.get("lr_decay", 0.995)),
                                              smooth=tr.get("smooth") if grad_smooth else None,
                                              seabed=seabed,
                                              source_illumination=source_illumination, log=log)
        return p, models, hist
    lr = float(lr["vp"] if isinstance(lr, dict) else lr)
    vp, hist = run_multiscale(p, scales, int(tr["N_epochs"]), lr=lr, optimizer=optimizer,
                              scale_decay=float(tr.get("scale_decay", 0.85)),
                              lr_decay=float(tr.get("lr_decay", 0.995)),
                              smooth=tr.get("smooth") if grad_smooth else None, seabed=seabed,
                              source_illumination=source_illumination, log=log)
    return p, vp, hist


# ---------------------------------------------------------- source encoding

def encode_shots(rng, wavelet, records, shots, split=0.2):
    """codingfwi.py:214-234 with roll() of utils.py:201-215: every chosen shot
    gets a random polarity (-1)**p, p in {1, 2}, and a random time shift tau
    in [0, split*nt); its wavelet and its records are shifted by tau (the
    first tau samples zero) and signed.  wavelet: (nt,) or (B, nt);
    records: (nt, B, Rs).  Returns (wavelets (n, nt), summed records
    (nt, Rs), [(shot, sign, tau)])."""
    nt = records.shape[0]
    n_shift = int(np.ceil(split * nt))
    wavs = torch.zeros((len(shots), nt), device=records.device, dtype=torch.float32)
    summed = torch.zeros((nt, records.shape[2]), device=records.device, dtype=torch.float32)
    codes = []
    for i, b in enumerate(shots):
        p = int(rng.integers(1, 3))
        tau = int(rng.integers(0, n_shift))
        sign = -1.0 if p % 2 else 1.0
        w = wavelet if wavelet.dim() == 1 else wavelet[b]
        wavs[i, tau:] = sign * w[:nt - tau]
        summed[tau:] += sign * records[:nt - tau, b]
        codes.append((int(b), sign, tau))
    return wavs, summed, codes


def run_encoded(p, scales, epochs, lr=5.0, optimizer="adam", batch=None, split=0.2,
                scale_decay=0.85, lr_decay=0.995, grad_clamp=True, smooth=None, seabed=None,
                seed=20230503, log=None):
    """The source-encoding FWI loop (codingfwi.py:184-334) on one GPU: per
    frequency scale the observed records and the wavelet are band-filtered
    (scipy backend, as codingfwi does, signal.py:49-96); per epoch all shots,
    or `batch` random ones, are encoded (encode_shots) into one field -- the
    first chosen shot's source carries its coded wavelet, the others enter as
    extra sources -- and one forward and one adjoint give the L2-misfit
    gradient against the summed coded records; post-processing and the update
    as in run_multiscale.  Needs every shot to record at the same receiver
    cells (the reference's fixed-receiver case).  Returns (vp, History)."""
    from . import signal as sg
    nt = p.wavelet.shape[0]
    nz, nx = p.vp_init.shape
    plane = nz * nx
    Rs = p.rec.numel() // p.B
    rf = p.rec.view(p.B, Rs) - torch.arange(p.B, device=p.rec.device)[:, None] * plane
    if not bool((rf == rf[0]).all()):
        raise ValueError("source encoding needs the same receiver cells for every shot")
    rec0 = rf[0].contiguous()
    rng = np.random.default_rng(seed)
    seis = sg.SeisSignal(dt=p.dt, filter_ord=p.filter_ord)
    obs_all = p.obs.view(nt, p.B, Rs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    vi = p.interior(p.vp_init).clone().contiguous()
    hist = History()
    for si, band in enumerate(scales):
        if band == "all":
            obs_b, wav_b = obs_all, p.wavelet
        else:
            host = obs_all.permute(1, 0, 2).cpu().numpy()
            filt = seis.filter(list(host), band, backend="scipy")
            obs_b = torch.from_numpy(np.stack(list(filt), 1)).to(p.obs.device)
            wav_b = torch.from_numpy(seis.filter([p.wavelet.cpu().numpy()], band,
                                                 backend="scipy")[0]).to(p.obs.device)
        slr = lr * scale_decay ** si
        x = torch.nn.Parameter(vi.clone())
        opt = torch.optim.Adam([x], lr=slr, eps=1e-22) if optimizer == "adam" else None
        sched = torch.optim.lr_scheduler.ExponentialLR(opt, lr_decay) if opt else None
        cur = slr
        for ep in range(epochs):
            shots = (rng.choice(p.B, batch, replace=False) if batch else np.arange(p.B)).tolist()
            wavs, coded_obs, _ = encode_shots(rng, wav_b, obs_b, shots, split)
            src = p.src[shots[0]:shots[0] + 1].contiguous()
            es = torch.zeros((len(shots) - 1, 3), dtype=torch.int32, device=src.device)
            if len(shots) > 1:
                es[:, 1:] = p.src[shots[1:]]
            v = p.pad(x.detach()).requires_grad_(True)
            syn = ac.propagate(v, p.d, wavs[0].contiguous(), src, rec0, 1, p.dt, p.h,
                               extra_sources=(es, wavs[1:].contiguous()))
            loss = torch.sum((syn - coded_obs) ** 2)
            loss.backward()
            g = v.grad
            if p.mute:
                g[PAD:PAD + p.mute] = 0.0
            g = postprocess_padded(g, smooth, seabed)
            if opt is not None:
                opt.zero_grad()
                x.grad = g.contiguous()
                opt.step()
                sched.step()
            else:
                with torch.no_grad():
                    x -= cur * _gd_direction(g, grad_clamp)
                cur *= lr_decay
            with torch.no_grad():
                x.clamp_(p.vmin, p.vmax)
            torch.cuda.synchronize()
            hist.add(time.perf_counter() - t0, float(loss), p.rmse(p.pad(x.detach())),
                     f"scale{si}")
            if log:
                log(f"[encoded] band {band} epoch {ep + 1}: shots {len(shots)}, misfit "
                    f"{float(loss):.4e} rmse {hist.rmse[-1]:.1f}")
        vi = x.detach().contiguous()
    return p.pad(vi), hist


# ---------------------------------------------------------- elastic FWI

ELASTIC_BOUNDS = {"vp": (1400.0, 6000.0), "vs": (0.0, 3500.0), "rho": (1000.0, 3000.0)}


@dataclass
class ElasticProblem:
    """Elastic FWI on one GPU (seistorch_dist.py:166-303 with
    equation: elastic): models_true / models_init hold padded (nz, nx)
    'vp', 'vs', 'rho'; invlist names the inverted ones (geom.invlist);
    records (nt, R, nch) of rec_type from sources on src_type."""
    models_true: dict
    models_init: dict
    d: torch.Tensor
    wavelet: torch.Tensor
    src: torch.Tensor
    rec: torch.Tensor
    B: int
    dt: float
    h: float
    src_type: tuple = ("txx", "tzz")
    rec_type: tuple = ("vx", "vz")
    invlist: tuple = ("vp", "vs")
    obs: torch.Tensor = None
    mute: int = 5
    band: object = "all"
    filter_ord: int = 3
    _fobs: tuple = None

    interior = Problem.interior
    pad = Problem.pad
    _filter = Problem._filter
    observed = Problem.observed

    def _propagate(self, m, illum=None):
        from .equations2d import elastic as el
        return el.propagate(m["vp"], m["vs"], m["rho"], self.d, self.wavelet, self.src, self.rec,
                            self.B, self.dt, self.h, src_type=self.src_type,
                            rec_type=self.rec_type, illum=illum)

    def records(self, m):
        with torch.no_grad():
            return self._propagate(m)

    def gradient(self, m, illum=None, padded=False):
        """(misfit, {name: gradient} for the inverted parameters): interior
        gradients, or over the padded grid with padded=True (see
        Problem.gradient); illum: optional (nz, nx) buffer for the forward's
        source illumination."""
        mm = {k: (v.detach().clone().requires_grad_(k in self.invlist)) for k, v in m.items()}
        loss = torch.sum((self._filter(self._propagate(mm, illum)) - self.observed()) ** 2)
        loss.backward()
        g = {}
        for k in self.invlist:
            gk = mm[k].grad.clone()
            if self.mute:
                gk[PAD:PAD + self.mute] = 0.0
            g[k] = gk if padded else self.interior(gk).clone()
        return float(loss), g

    def rmse(self, m, name="vp"):
        return float(torch.sqrt(torch.mean((self.interior(m[name])
                                            - self.interior(self.models_true[name])) ** 2)))


@dataclass
class RhoProblem(ElasticProblem):
    """Variable-density acoustic FWI (equation: acoustic_rho_habc,
    eqconfigure.py:11, acoustic_rho_habc.py:31-121): models_true /
    models_init hold padded (nz, nx) 'vp' and 'rho', d is the HABC weight
    field, records (nt, R) of the pressure; invlist names the inverted
    parameters (geom.invlist, default both).  The gradients come from the
    boundary-saving adjoint of seis_rho_habc_backward; run_multiscale_elastic
    drives the updates (one learning rate per parameter, ELASTIC_BOUNDS)."""
    src_type: tuple = ("p",)
    rec_type: tuple = ("p",)
    invlist: tuple = ("vp", "rho")

    def _propagate(self, m, illum=None):
        from .equations2d import acoustic_rho_habc as ar
        return ar.propagate(m["vp"], m["rho"], self.d, self.wavelet, self.src, self.rec, self.B,
                            self.dt, self.h, illum=illum)


def run_multiscale_elastic(p, scales, epochs, lr, optimizer="adam", scale_decay=0.85,
                           lr_decay=0.995, grad_clamp=True, smooth=None, seabed=None,
                           source_illumination=False, log=None):
    """run_multiscale for the elastic equation: one fresh optimiser per scale
    over the inverted parameters, each with its own learning rate (lr: a
    number or {name: lr}, the reference's --lr vp=.. vs=..), scaled by
    scale_decay**scale and decayed by lr_decay per epoch; Adam (eps 1e-22) or
    max-normalised steepest descent per parameter; each parameter clamped to
    ELASTIC_BOUNDS after the update; smooth, seabed and source_illumination
    post-process every inverted parameter's gradient as in run_multiscale.
    Returns (models, History)."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    inv = list(p.invlist)
    lrs = {k: float(lr[k] if isinstance(lr, dict) else lr) for k in inv}
    cur_int = {k: p.interior(p.models_init[k]).clone().contiguous() for k in inv}
    hist = History()
    band0 = p.band

    def models_of(xs):
        m = dict(p.models_init)
        for k in inv:
            m[k] = p.pad(xs[k].detach())
        return m

    try:
        for si, band in enumerate(scales):
            p.band = band
            xs = {k: torch.nn.Parameter(cur_int[k].clone()) for k in inv}
            groups = [{"params": [xs[k]], "lr": lrs[k] * scale_decay ** si} for k in inv]
            opt = torch.optim.Adam(groups, eps=1e-22) if optimizer == "adam" else None
            sched = torch.optim.lr_scheduler.ExponentialLR(opt, lr_decay) if opt else None
            step = {k: lrs[k] * scale_decay ** si for k in inv}
            for ep in range(epochs):
                illum = torch.empty_like(p.models_init["vp"]) if source_illumination else None
                loss, g = p.gradient(models_of(xs), illum, padded=True)
                for k in inv:
                    g[k] = postprocess_padded(g[k], smooth, seabed, illum)
                if opt is not None:
                    opt.zero_grad()
                    for k in inv:
                        xs[k].grad = g[k].contiguous()
                    opt.step()
                    sched.step()
                else:
                    with torch.no_grad():
                        for k in inv:
                            xs[k] -= step[k] * _gd_direction(g[k], grad_clamp)
                            step[k] *= lr_decay
                with torch.no_grad():
                    for k in inv:
                        xs[k].clamp_(*ELASTIC_BOUNDS[k])
                torch.cuda.synchronize()
                hist.add(time.perf_counter() - t0, loss, p.rmse(models_of(xs)), f"scale{si}")
                if log:
                    log(f"[elastic multiscale] band {band} epoch {ep + 1}: misfit {loss:.4e} "
                        f"vp rmse {hist.rmse[-1]:.1f}")
            cur_int = {k: xs[k].detach().contiguous() for k in inv}
    finally:
        p.band = band0
    return models_of(cur_int), hist
