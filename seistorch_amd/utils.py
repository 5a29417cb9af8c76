"""Small host utilities with the reference's semantics."""
import numpy as np
import torch


def ricker_wave(fm, dt, T, delay=80, dtype='tensor', inverse=False):
    """Ricker wavelet, seistorch/utils.py:235-249 (delay in samples, FP64 -> FP32)."""
    i = np.arange(T, dtype=np.float64)
    c = np.pi * fm * (i * dt - delay * dt)
    p = -1.0 if inverse else 1.0
    r = (p * (1 - 2 * np.power(c, 2)) * np.exp(-np.power(c, 2))).astype(np.float32)
    return r if dtype == 'numpy' else torch.from_numpy(r)


def to_tensor(x, dtype=None):
    dtype = dtype if dtype is not None else torch.get_default_dtype()
    if torch.is_tensor(x):
        return x.to(dtype)
    return torch.as_tensor(np.asarray(x)).to(dtype)


def pad_model(m, bwidth=50, multiple=False):
    """Edge padding of a (nz, nx) model, seistorch/geom.py:213-240."""
    top = 0 if multiple else bwidth
    return np.pad(np.asarray(m), [[top, bwidth], [bwidth, bwidth]], mode="edge")
