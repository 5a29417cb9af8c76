"""ctypes wrapper of oracle/fd64.c: the FP64 acoustic PML oracle in C, for
parity checks at full configuration sizes.

TEST INFRASTRUCTURE: imported only by tests/; never by the product.

Same inputs and outputs as ``oracle.acoustic.forward`` / ``gradient`` except
that records come back as one (nt, R) array (the shots' receiver columns side
by side, the layout of the HIP ``propagate``), and the gradient takes the
records' cotangent instead of observed records, so it serves any functional
of the records (the L2 misfit's cotangent is 2*(syn - obs),
seistorch/loss.py:409-421).  Pinned against ``oracle.acoustic`` in
tests/test_oracle_fd64.py.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfd64.so")
_lib = None


def build():
    """Compile libfd64.so with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", os.path.dirname(_HERE), "-f",
                    os.path.join(_HERE, "Makefile")], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        p, i = ctypes.c_void_p, ctypes.c_int64
        dbl = ctypes.c_double
        lib.fd64_acoustic_forward.argtypes = [p, p, i, i, dbl, dbl, p, i, p, i, p, p, p]
        lib.fd64_acoustic_gradient.argtypes = [p, p, i, i, dbl, dbl, p, i, p, i, p, p, p, i, p]
        lib.fd64_acoustic_forward.restype = ctypes.c_int
        lib.fd64_acoustic_gradient.restype = ctypes.c_int
        _lib = lib
    return _lib


def _prep(c, d, wavelet, src, rec):
    c = np.ascontiguousarray(c, np.float64)
    d = np.ascontiguousarray(d, np.float64)
    w = np.ascontiguousarray(wavelet, np.float64).ravel()
    src = np.ascontiguousarray(np.asarray(src, np.int64).reshape(-1, 2))
    recs = [np.asarray(r, np.int64).reshape(-1, 2) for r in rec]
    assert len(recs) == src.shape[0]
    nz, nx = c.shape
    for r in [src] + recs:
        assert ((r[:, 0] >= 0) & (r[:, 0] < nz) & (r[:, 1] >= 0) & (r[:, 1] < nx)).all()
    off = np.zeros(len(recs) + 1, np.int64)
    off[1:] = np.cumsum([len(r) for r in recs])
    rz = np.ascontiguousarray(np.concatenate(recs, 0) if recs else np.zeros((0, 2), np.int64))
    return c, d, w, src, rz, off


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def forward(c, d, dt, h, wavelet, src, rec):
    """Records (nt, R) in FP64 of the shots (src (B, 2), rec: B arrays (n_b, 2))."""
    c, d, w, src, rz, off = _prep(c, d, wavelet, src, rec)
    nz, nx = c.shape
    out = np.zeros((w.size, int(off[-1])), np.float64)
    rc = _load().fd64_acoustic_forward(_p(c), _p(d), nz, nx, float(dt), float(h), _p(w), w.size,
                                       _p(src), src.shape[0], _p(rz), _p(off), _p(out))
    if rc != 0:
        raise MemoryError("fd64_acoustic_forward: out of memory")
    return out


def gradient(c, d, dt, h, wavelet, src, rec, rec_cot, every=100):
    """dJ/dc (nz, nx) in FP64 for dJ/d records = rec_cot (nt, R)."""
    c, d, w, src, rz, off = _prep(c, d, wavelet, src, rec)
    nz, nx = c.shape
    cot = np.ascontiguousarray(rec_cot, np.float64)
    assert cot.shape == (w.size, int(off[-1]))
    g = np.zeros((nz, nx), np.float64)
    rc = _load().fd64_acoustic_gradient(_p(c), _p(d), nz, nx, float(dt), float(h), _p(w), w.size,
                                        _p(src), src.shape[0], _p(rz), _p(off), _p(cot),
                                        int(every), _p(g))
    if rc != 0:
        raise MemoryError("fd64_acoustic_gradient: out of memory")
    return g
