"""Record filtering for multiscale FWI on the HIP kernel ``seis_filtfilt``.

Mirrors seistorch/signal.py:11-96 (``SeisSignal.decide_filter_type`` and
``SeisSignal.filter``): Butterworth (b, a) from scipy.signal.butter with
Wn = 2 f dt; backend 'torch' (the FWI loop's filter of synthetic and
observed records, seistorch_dist.py:222-225) runs the zero-phase filter of
torchaudio.functional.filtfilt(x.double(), a, b, clamp=False).float() along
time on the GPU: zero-state recursion forward, then backward, in FP64, no
edge padding.  It is differentiable; the filter is self-adjoint, so its
backward is the same kernel on the cotangent.  Backend 'scipy' is the
reference's CPU path for numpy arrays (scipy.signal.filtfilt, padded), which
the FWI loop uses on the 1-D source wavelet (seistorch_dist.py:179-182).
"""
import ctypes

import numpy as np
import torch

from . import _lib

VALID_MODES = ("lowpass", "highpass", "bandpass")


def decide_filter_type(freq):
    """seistorch/signal.py:25-37."""
    if isinstance(freq, str):
        if freq == "all":
            return "all"
        raise ValueError(f"unknown frequency spec {freq!r}")
    if isinstance(freq, (int, float)):
        return "lowpass"
    if isinstance(freq, (list, tuple)):
        if len(freq) == 1:
            return "lowpass"
        if len(freq) == 2:
            return "bandpass"
    raise ValueError(f"unsupported frequency spec {freq!r}")


def butter(order, freqs, dt):
    """(b, a) FP64 of seistorch/signal.py:59-75 (scipy.signal.butter)."""
    from scipy import signal
    mode = decide_filter_type(freqs)
    assert mode in VALID_MODES, "mode must be lowpass, highpass or bandpass"
    if mode == "lowpass" and isinstance(freqs, (list, tuple)):
        freqs = freqs[0]
    f = [freqs] if isinstance(freqs, (int, float)) else list(freqs)
    wn = [2 * fr / (1 / dt) for fr in f]
    wn = wn[0] if len(wn) == 1 else wn
    b, a = signal.butter(order, Wn=wn, btype=mode)
    return np.ascontiguousarray(b, np.float64), np.ascontiguousarray(a, np.float64)


def _run(x, b, a):
    if not x.is_cuda or x.dtype != torch.float32:
        raise RuntimeError("filtfilt needs a float32 tensor on the HIP device")
    x = x.contiguous()
    nt = x.shape[0]
    ntr = x.numel() // max(nt, 1)
    y = torch.empty_like(x)
    if x.numel() == 0:
        return y
    work = torch.empty((nt, ntr), dtype=torch.float64, device=x.device)
    bb = np.ascontiguousarray(b, np.float64)
    aa = np.ascontiguousarray(a, np.float64)
    if bb.shape != aa.shape:
        n = max(bb.size, aa.size)
        bb = np.pad(bb, (0, n - bb.size))
        aa = np.pad(aa, (0, n - aa.size))
    _lib.call("seis_filtfilt", _lib.ptr(x), _lib.ptr(y), _lib.ptr(work), nt, ntr,
              bb.ctypes.data_as(ctypes.c_void_p), aa.ctypes.data_as(ctypes.c_void_p),
              int(bb.size), _lib.stream_of(x))
    return y


class _FiltFilt(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, b, a):
        ctx.ba = (b, a)
        return _run(x, b, a)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gy):
        b, a = ctx.ba
        return _run(gy.contiguous().to(torch.float32), b, a), None, None


def filtfilt(x, b, a):
    """Zero-phase filter of x (nt, ...) along dim 0 (time); differentiable."""
    return _FiltFilt.apply(x, np.asarray(b, np.float64), np.asarray(a, np.float64))


class SeisSignal:
    """seistorch/signal.py:11-23 with the configuration reduced to what the
    filter needs: ``dt`` (cfg['geom']['dt']) and ``filter_ord``
    (cfg['training']['filter_ord'])."""

    def __init__(self, cfg=None, logger=None, dt=None, filter_ord=None):
        self.cfg = cfg
        self.logger = logger
        self.dt = cfg["geom"]["dt"] if cfg is not None else dt
        self.forder = cfg["training"]["filter_ord"] if cfg is not None else filter_ord

    def decide_filter_type(self, freq):
        return decide_filter_type(freq)

    def filter(self, d, freqs, axis=0, threads=1, backend="scipy", **kwargs):
        """Filter records (seistorch/signal.py:49-96).  backend 'torch': d is a
        (nt, ...) tensor or a list of them (one per shot), filtered along time
        on the GPU, returned as new tensors; backend 'scipy': d is an array of
        per-shot numpy arrays, filtered with scipy.signal.filtfilt on axis."""
        mode = decide_filter_type(freqs)
        if self.logger is not None:
            self.logger.print(f"Data filtering (mode: {mode}): frequency:{freqs}")
        if freqs == "all":
            return d
        b, a = butter(self.forder, freqs, self.dt)
        if backend == "torch":
            if isinstance(d, (list, tuple)):
                return type(d)(filtfilt(x, b, a) for x in d)
            return filtfilt(d, b, a)
        if backend == "scipy":
            from scipy import signal
            out = np.empty(len(d), dtype=object)
            out[:] = [signal.filtfilt(b, a, di, axis=axis).astype(np.float32) for di in d]
            return out
        raise ValueError(f"unknown backend {backend!r}")
