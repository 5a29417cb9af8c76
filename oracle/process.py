"""FP64 restatement of seistorch's gradient post-processing.

TEST INFRASTRUCTURE: imported only by tests/, smoke() and bench.py's
cpu_baseline leg.

seistorch/signal.py:247-304 (gaussian_filter, 2-D): taps
w = exp(-(k - ks//2)^2 / (2 sigma^2)) / sum, ks = 2r+1 for an even radius r
and 2r for an odd one (:263-266); the field is reflect-padded (numpy
'reflect') by ks//2 on both axes, correlated along one axis with zero padding
ks//2 (F.conv2d) and cropped by ks//2 on both axes.  Written out:
    out[i] = sum_k w[k] x[refl(i + k - ks//2)],  i < n + 2(ks//2) - ks + 1.
seistorch/process.py: smooth_gradient (:66-112: grad * mask when the seabed
cut is on, then `counts` z and x passes), cut_gradient (:43-47: grad * mask),
precondition (:114-117: grad / illumination), load_seabed (:26-41: zero
padding by the boundary width, no top padding with multiples), applied in
the order of seistorch_dist.py:270-280.

Pinned by tests/test_oracle_process.py: the closed form against the
reference's own steps (reflect pad -> zero-padded correlation -> crop) run
with numpy's pad and torch's conv2d, for odd and even kernel sizes on both
axes; against scipy.ndimage.correlate1d (mode 'mirror' = numpy 'reflect')
for odd sizes; constants are preserved (taps sum to 1); a centred impulse
far from the edges reproduces the taps; the FP32 taps match the reference
formula's values.
"""
import numpy as np


def kernel_size(radius):
    """signal.py:263-266."""
    return 2 * radius + 1 if radius % 2 == 0 else 2 * radius


def taps64(sigma, radius):
    ks = kernel_size(radius)
    w = np.exp(-(np.arange(ks) - ks // 2) ** 2 / (2 * sigma ** 2))
    return w / w.sum()


def _reflect(j, n):
    j = np.abs(j)
    return np.where(j >= n, 2 * (n - 1) - j, j)


def gaussian_filter_taps(x, w, axis):
    """One pass with the given taps (any float dtype), FP64 accumulation."""
    x = np.asarray(x, np.float64)
    ks = len(w)
    r = ks // 2
    if axis == 1:
        return gaussian_filter_taps(x.T, w, 0).T
    n = x.shape[0]
    nout = n + 2 * r - ks + 1
    i = np.arange(nout)
    out = np.zeros((nout, x.shape[1]))
    for k in range(ks):
        out += float(w[k]) * x[_reflect(i + k - r, n)]
    return out


def gaussian_filter(x, sigma, radius, axis):
    """signal.py:247-304 with FP64 taps."""
    return gaussian_filter_taps(x, taps64(sigma, radius), axis)


def postprocess(grad, smooth=None, pre_mask=None, post_mask=None, illum=None, taps=None):
    """seistorch_dist.py:270-280 on one gradient.  taps: optional (wz, wx)
    to use instead of the FP64 ones (e.g. the FP32 taps the GPU uses)."""
    g = np.asarray(grad, np.float64)
    if pre_mask is not None:
        g = g * pre_mask
    if smooth is not None:
        if taps is None:
            wz = taps64(smooth["sigma"]["z"], smooth["radius"]["z"])
            wx = taps64(smooth["sigma"]["x"], smooth["radius"]["x"])
        else:
            wz, wx = taps
        for _ in range(int(smooth["counts"])):
            g = gaussian_filter_taps(g, wz, 0)
            g = gaussian_filter_taps(g, wx, 1)
    if post_mask is not None:
        g = g * post_mask
    if illum is not None:
        g = g / illum
    return g


def seabed_mask(seabed, padding, multiple):
    """process.py:26-41."""
    top = 0 if multiple else padding
    return np.pad(np.asarray(seabed, np.float64), ((top, padding), (padding, padding)))
