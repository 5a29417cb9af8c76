"""CPU pinning of the record-filter oracle (oracle/signal.py)."""
import numpy as np
import pytest
from scipy import signal as ss

from oracle import signal as osg


def _ba(order=3, freqs=5.0, dt=1e-3):
    return osg.butter_ba(order, freqs, dt)


# (freqs, dt, tolerance relative to the peak): the direct-form recursion of a
# Butterworth (b, a) pair loses accuracy as Wn = 2 f dt shrinks (the a
# coefficients nearly cancel), so two summation orders agree only to the
# filter's conditioning; the 1-3 Hz band at dt = 1 ms is the reference's
# template setting (template.yml:44,102-105)
@pytest.mark.parametrize("freqs,dt,tol", [(5.0, 1e-3, 1e-9), ([5.0], 4e-3, 1e-11),
                                          ([4.0, 12.0], 4e-3, 1e-9),
                                          ([1.0, 3.0], 1e-3, 5e-5)])
def test_lfilter_matches_scipy_zero_state(freqs, dt, tol):
    b, a = _ba(3, freqs, dt)
    rng = np.random.default_rng(0)
    x = rng.standard_normal((300, 5))
    ref = ss.lfilter(b, a, x, axis=0)
    got = osg.lfilter(b, a, x)
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()


def test_lfilter_brute_force_convolution():
    """y = h * x with h the impulse response of the recursion (FIR check of
    the whole recursion on a tiny case, scipy-independent)."""
    b, a = np.array([0.3, 0.2, -0.1]), np.array([1.0, -0.5, 0.25])
    h = np.zeros(40)
    for t in range(40):  # impulse response from the difference equation
        h[t] = (b[t] if t < 3 else 0.0) - (0.0 if t < 1 else a[1] * h[t - 1]) \
            - (0.0 if t < 2 else a[2] * h[t - 2])
    x = np.random.default_rng(3).standard_normal((40, 2))
    ref = np.stack([np.convolve(h, x[:, j])[:40] for j in range(2)], 1)
    assert np.abs(osg.lfilter(b, a, x) - ref).max() <= 1e-12


def test_filtfilt_is_self_adjoint():
    b, a = _ba(3, [4.0, 16.0], 4e-3)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((400, 3))
    y = rng.standard_normal((400, 3))
    lhs = np.sum(osg.filtfilt(b, a, x) * y)
    rhs = np.sum(x * osg.filtfilt(b, a, y))
    assert abs(lhs - rhs) <= 1e-10 * (abs(lhs) + abs(rhs))


def test_filtfilt_zero_phase_impulse():
    b, a = _ba(3, 20.0)
    x = np.zeros((801, 1))
    x[400] = 1.0
    h = osg.filtfilt(b, a, x)[:, 0]
    k = np.arange(1, 300)
    assert np.abs(h[400 + k] - h[400 - k]).max() <= 1e-9 * np.abs(h).max()
    assert h.argmax() == 400


@pytest.mark.parametrize("f", [2.0, 5.0, 9.0])
def test_filtfilt_gain_on_sinusoid(f):
    dt = 1e-3
    b, a = _ba(3, 5.0, dt)
    t = np.arange(6000) * dt
    x = np.sin(2 * np.pi * f * t)[:, None]
    y = osg.filtfilt(b, a, x)[:, 0]
    _, H = ss.freqz(b, a, worN=[f], fs=1.0 / dt)
    mid = slice(2500, 3500)
    amp = np.sqrt(2.0 * np.mean(y[mid] ** 2))
    assert abs(amp - abs(H[0]) ** 2) <= 2e-3


def test_frequency_spec_modes():
    assert osg.filter_mode(5) == "lowpass"
    assert osg.filter_mode([5.0]) == "lowpass"
    assert osg.filter_mode([1.0, 3.0]) == "bandpass"
    assert osg.filter_mode("all") == "all"
    b, a = _ba(3, [1.0, 3.0], 1e-3)
    assert len(b) == len(a) == 7      # bandpass doubles the order
    b, a = _ba(3, 5.0, 1e-3)
    assert len(b) == len(a) == 4
    rb, ra = ss.butter(3, 2 * 5.0 * 1e-3, btype="lowpass")
    assert np.array_equal(b, rb) and np.array_equal(a, ra)


def test_all_passes_through():
    x = np.random.default_rng(2).standard_normal((10, 4)).astype(np.float32)
    assert np.array_equal(osg.filter_records(x, "all", 1e-3, 3), x)
