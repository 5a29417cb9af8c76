"""Gradient post-processing on the HIP kernels ``seis_gaussian_filter`` and
``seis_smooth_gradient``.

Mirrors seistorch/process.py:9-117 (``PostProcess``: ``load_seabed``,
``cut_gradient``, ``smooth_gradient``, ``precondition``) and the 2-D case of
seistorch/signal.py:247-304 (``gaussian_filter``), applied by the FWI loop
between ``loss.backward()`` and the optimizer step in the order
smooth -> cut -> precondition (seistorch_dist.py:270-280).  The reference
smooths on the CPU (the gradient goes through numpy for the reflect
padding); here every step runs on the gradient in HBM.
"""
import ctypes

import numpy as np
import torch

from . import _lib


def gaussian_kernel(sigma, radius):
    """The taps of seistorch/signal.py:262-275: ks = 2r+1 for an even radius,
    2r for an odd one, exp(-(k - ks//2)^2 / (2 sigma^2)) in FP64, cast to FP32
    and normalised to sum 1 in FP32."""
    radius = int(radius)
    if radius < 0:
        raise ValueError(f"radius must be >= 0, got {radius}")
    ks = 2 * radius + 1 if radius % 2 == 0 else 2 * radius
    k = torch.tensor(np.exp(-(np.arange(ks) - ks // 2) ** 2 / (2 * sigma ** 2)),
                     dtype=torch.float32)
    return (k / k.sum()).numpy()


def _host_taps(w):
    w = np.ascontiguousarray(w, np.float32)
    return w, w.ctypes.data_as(ctypes.c_void_p)


def _check_field(t, name):
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32:
        raise RuntimeError(f"{name} must be a float32 tensor on the HIP device")
    if t.ndim != 2:
        raise ValueError(f"{name} must be 2-D (nz, nx), got shape {tuple(t.shape)}")
    return t.contiguous()


def gaussian_filter(input_tensor, sigma, radius, axis=-1):
    """One 1-D Gaussian pass over a 2-D tensor (seistorch/signal.py:247-304).

    The reference reflect-pads by ks//2, correlates with zero padding and
    crops; the result is
    out[i] = sum_k w[k] x[reflect(i + k - ks//2)] along ``axis``, with one
    more cell than the input along it when the radius is odd (even ks).
    Axis -1 / -2 are taken as 1 / 0 (the reference accepts only 0 and 1)."""
    x = _check_field(input_tensor, "input_tensor")
    if axis not in (0, 1, -1, -2):
        raise ValueError(f"axis must be 0 or 1, got {axis}")
    axis = axis % 2
    w, wp = _host_taps(gaussian_kernel(sigma, radius))
    ks = w.size
    nz, nx = x.shape
    grow = 2 * (ks // 2) - ks + 1
    shape = (nz + grow, nx) if axis == 0 else (nz, nx + grow)
    y = torch.empty(shape, dtype=torch.float32, device=x.device)
    _lib.call("seis_gaussian_filter", _lib.ptr(x), _lib.ptr(y), nz, nx, axis, wp, ks,
              _lib.stream_of(x))
    return y


def smooth_gradient(grad, smooth=None, pre_mask=None, post_mask=None, illum=None):
    """The whole post-processing of one gradient (nz, nx) in one call:
    grad * pre_mask, then ``smooth['counts']`` z and x Gaussian passes
    (``smooth`` = cfg['training']['smooth'], with sigma and radius per axis
    'z' and 'x'), then * post_mask, then / illum; each step only when its
    argument is given.  Returns a new tensor."""
    g = _check_field(grad, "grad")
    nz, nx = g.shape
    masks = []
    for name, m in (("pre_mask", pre_mask), ("post_mask", post_mask), ("illum", illum)):
        if m is not None:
            m = _check_field(m, name)
            if m.shape != g.shape:
                raise ValueError(f"{name} shape {tuple(m.shape)} != gradient shape {tuple(g.shape)}")
        masks.append(m)
    counts = int(smooth["counts"]) if smooth is not None else 0
    out = torch.empty_like(g)
    work = torch.empty_like(g) if counts > 0 else None
    wz, wzp = _host_taps(gaussian_kernel(smooth["sigma"]["z"], smooth["radius"]["z"])
                         if counts > 0 else np.ones(1, np.float32))
    wx, wxp = _host_taps(gaussian_kernel(smooth["sigma"]["x"], smooth["radius"]["x"])
                         if counts > 0 else np.ones(1, np.float32))
    _lib.call("seis_smooth_gradient", _lib.ptr(g), _lib.ptr(out),
              _lib.ptr(work) if work is not None else None, nz, nx, wzp, wz.size, wxp, wx.size,
              counts, *[_lib.ptr(m) if m is not None else None for m in masks],
              _lib.stream_of(g))
    return out


class PostProcess:
    """seistorch/process.py:7-117.  ``model`` is anything with
    ``parameters()`` (an ``nn.Module`` or a list of tensors wrapped by
    ``Params``); its ``precondition`` attribute, when present, is the source
    illumination (sum over steps and shots of the squared source-side field,
    rnn.py:204-205).  ``cfg`` holds geom.boundary.width, geom.multiple,
    geom.seabed (a .npy path, loaded without pickles) and
    training.smooth; ``commands`` has ``grad_cut``."""

    def __init__(self, model, cfg, commands):
        self.model = model
        self.cfg = cfg
        self.commands = commands
        self.modelmask = self.load_seabed() if getattr(commands, "grad_cut", False) else None

    def load_seabed(self):
        """process.py:26-41: the seabed mask padded with zeros by the boundary
        width (no top padding when multiples are modelled)."""
        padding = self.cfg["geom"]["boundary"]["width"]
        multiple = self.cfg["geom"]["multiple"]
        seabed = self.cfg["geom"]["seabed"]
        if isinstance(seabed, str):
            seabed = np.load(seabed, allow_pickle=False)
        seabed = torch.as_tensor(np.asarray(seabed), dtype=torch.float32)
        top = 0 if multiple else padding
        return torch.nn.functional.pad(seabed, (padding, padding, top, padding), mode="constant",
                                       value=0)

    def _params(self):
        return [p for p in self.model.parameters() if p.requires_grad]

    def _mask_on(self, p):
        if self.modelmask.device != p.device:
            self.modelmask = self.modelmask.to(p.device)
        return self.modelmask

    def cut_gradient(self):
        """process.py:43-47: grad *= mask."""
        for p in self._params():
            p.grad = smooth_gradient(p.grad, post_mask=self._mask_on(p))

    def smooth_gradient(self):
        """process.py:66-112: grad *= mask when the seabed cut is on, then
        ``counts`` z and x Gaussian passes."""
        smooth = self.cfg["training"]["smooth"]
        for p in self._params():
            pre = self._mask_on(p) if self.modelmask is not None else None
            p.grad = smooth_gradient(p.grad, smooth, pre_mask=pre)

    def precondition(self):
        """process.py:114-117: grad /= the source illumination."""
        illum = self.model.precondition
        for p in self._params():
            p.grad = smooth_gradient(p.grad, illum=illum.to(p.device))

    def apply(self, grad_smooth=False, grad_cut=False, source_illumination=False):
        """All enabled steps of seistorch_dist.py:270-280 in one kernel
        sequence per parameter (same results as calling the methods in turn)."""
        smooth = self.cfg["training"]["smooth"] if grad_smooth else None
        for p in self._params():
            mask = self._mask_on(p) if grad_cut else None
            p.grad = smooth_gradient(
                p.grad, smooth, pre_mask=mask if grad_smooth else None, post_mask=mask,
                illum=self.model.precondition.to(p.device) if source_illumination else None)


class Params:
    """A list of tensors seen as a model: ``parameters()`` and an optional
    ``precondition`` (illumination) attribute."""

    def __init__(self, tensors, precondition=None):
        self._t = list(tensors)
        self.precondition = precondition

    def parameters(self):
        return iter(self._t)
