"""The C FP64 oracle (oracle/fd64.c) equals the numpy FP64 oracle
(oracle/acoustic.py, pinned by closed forms in test_oracle_acoustic.py) on
small cases: ragged and repeated receivers, a receiver on the source cell,
checkpoint intervals that do and do not divide nt.  Both are FP64 with
different summation orders, so they agree to ~1e-13."""
import numpy as np
import pytest

from oracle import acoustic as oa
from oracle import fd64
from tests.cases import acoustic_case


def _ragged(c):
    recs = [c.rec[0], c.rec[1][::3], np.vstack([c.rec[2][:5], c.rec[2][:5], c.src[2:3]])]
    return recs


def test_forward_matches_numpy_oracle():
    c = acoustic_case(nz_i=20, nx_i=30, nshots=3, nt=150)
    recs = _ragged(c)
    ro = oa.forward(c.vp, c.d, c.dt, c.h, c.w, c.src, recs)
    ref = np.concatenate([r[..., 0] for r in ro], 1)
    got = fd64.forward(c.vp, c.d, c.dt, c.h, c.w, c.src, recs)
    assert got.shape == ref.shape
    e = np.abs(got - ref).max() / np.abs(ref).max()
    assert e < 1e-12, e


@pytest.mark.parametrize("every", [1, 7, 50, 150, 400])
def test_gradient_matches_numpy_oracle(every):
    c = acoustic_case(nz_i=16, nx_i=24, nshots=3, nt=150)
    recs = _ragged(c)
    rng = np.random.default_rng(every)
    obs = [rng.standard_normal((c.nt, len(r), 1)) * 1e-3 for r in recs]
    J, go, syn = oa.gradient(c.vp, c.d, c.dt, c.h, c.w, c.src, recs, obs)
    cot = 2.0 * np.concatenate([(s - o)[..., 0] for s, o in zip(syn, obs)], 1)
    g = fd64.gradient(c.vp, c.d, c.dt, c.h, c.w, c.src, recs, cot, every=every)
    e = np.abs(g - go).max() / np.abs(go).max()
    assert e < 1e-11, e


def test_gradient_of_linear_functional_is_the_born_transpose():
    """<G, (dr/dc) dc> = <(dr/dc)^T G, dc> with the numpy Born operator."""
    c = acoustic_case(nz_i=12, nx_i=18, nshots=2, nt=90)
    rng = np.random.default_rng(3)
    dc = rng.standard_normal(c.vp.shape)
    dr = np.concatenate([r[..., 0] for r in oa.born_apply(c.vp, c.d, c.dt, c.h, c.w, c.src,
                                                          c.rec, dc)], 1)
    G = rng.standard_normal(dr.shape)
    g = fd64.gradient(c.vp, c.d, c.dt, c.h, c.w, c.src, c.rec, G, every=13)
    lhs, rhs = float(np.sum(G * dr)), float(np.sum(g * dc))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs), (lhs, rhs)
