"""seis_smooth_gradient_noalloc (the post-processing entry point that keeps its
work plane in a static device workspace instead of an argument) against
seis_smooth_gradient with a caller-allocated work buffer, through the C-ABI.

Both run the same kernels in the same order, so the outputs are compared bit
for bit.  Two calls in a row, the second on a smaller grid, check that the
workspace's leftover contents from the first call do not leak into the
second; a grid larger than the workspace is refused before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from seistorch_amd import _lib
from seistorch_amd.process import gaussian_kernel

pytestmark = pytest.mark.gpu

MAX_CELLS = 1 << 22  # SEIS_NOALLOC_MAX_CELLS


def _taps(sigma, radius):
    w = np.ascontiguousarray(gaussian_kernel(sigma, radius), np.float32)
    return w, w.ctypes.data_as(ctypes.c_void_p), int(w.size)


def _inputs(nz, nx, seed):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((nz, nx)).astype(np.float32)
    mask = np.ones((nz, nx), np.float32)
    mask[: nz // 4] = 0.0
    illum = (rng.random((nz, nx)) + 0.1).astype(np.float32)
    d = lambda a: torch.from_numpy(a).cuda()
    return d(g), d(mask), d(illum)


def _run_both(nz, nx, counts, rz, rx, seed):
    g, mask, illum = _inputs(nz, nx, seed)
    wz, pz, kz = _taps(1.5, rz)
    wx, px, kx = _taps(2.5, rx)
    out_ref = torch.full_like(g, float("nan"))
    work = torch.full_like(g, float("nan"))
    out_na = torch.full_like(g, float("nan"))
    s = _lib.stream_of(g)
    P = _lib.ptr
    print(f"no-alloc smooth_gradient: grid {nz}x{nx}, counts {counts}, kz {kz}, kx {kx}",
          flush=True)
    _lib.call("seis_smooth_gradient", P(g), P(out_ref), P(work), nz, nx, pz, kz, px, kx, counts,
              P(mask), P(mask), P(illum), s)
    _lib.call("seis_smooth_gradient_noalloc", P(g), P(out_na), nz, nx, pz, kz, px, kx, counts,
              P(mask), P(mask), P(illum), s)
    torch.cuda.synchronize()
    return out_ref.cpu().numpy(), out_na.cpu().numpy()


def test_noalloc_equals_allocating_entry_point_twice():
    # first call on a C2-sized padded gradient (900 x 2500), then a smaller one
    for nz, nx, counts, rz, rx, seed in [(900, 2500, 2, 4, 6, 1), (61, 97, 3, 2, 2, 2)]:
        ref, got = _run_both(nz, nx, counts, rz, rx, seed)
        assert np.isfinite(ref).all()
        assert np.array_equal(ref, got), float(np.abs(ref - got).max())


def test_noalloc_rejects_grids_beyond_the_workspace():
    g = torch.zeros(8, device="cuda")  # never read: the size check fails first
    nz, nx = 2049, 2048  # one row more than 2^22 cells
    assert nz * nx > MAX_CELLS
    w = np.ones(1, np.float32)
    pw = w.ctypes.data_as(ctypes.c_void_p)
    with pytest.raises(RuntimeError, match="static workspace"):
        _lib.call("seis_smooth_gradient_noalloc", _lib.ptr(g), _lib.ptr(g), nz, nx, pw, 1, pw, 1,
                  1, None, None, None, _lib.stream_of(g))
