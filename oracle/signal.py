"""FP64 restatement of seistorch's record filtering for multiscale FWI.

TEST INFRASTRUCTURE: imported only by tests/, smoke() and bench.py's
cpu_baseline leg.

seistorch/signal.py:67-96: Butterworth (b, a) = scipy.signal.butter(order,
Wn = 2*f*dt, btype) with btype from the frequency spec (:25-37, :59-72);
backend 'torch' (the one the FWI loop uses on records,
seistorch_dist.py:222-225) applies torchaudio.functional.filtfilt(
x.double(), a, b, clamp=False) along time and casts back to float32.
torchaudio (absent here; its published algorithm) computes filtfilt as
lfilter(x) forward, then lfilter of the time-reversed result, reversed back,
where lfilter is the zero-initial-state direct-form recursion
    y[t] = (sum_k b[k] x[t-k]) / a[0] - sum_{k>=1} (a[k]/a[0]) y[t-k]
and no edge padding is applied (unlike scipy.signal.filtfilt).

Pinned by tests/test_oracle_signal.py: lfilter against scipy.signal.lfilter
(zero initial state, the textbook routine; to 1e-9 of the peak for a
well-conditioned filter, to the recursion's own conditioning, ~1e-5, for the
reference's 1-3 Hz band at dt = 1 ms), filtfilt's self-adjointness
(dot-product test), its zero phase (an impulse far from the edges gives a
symmetric response), its gain |H(f)|^2 on a sinusoid against
scipy.signal.freqz, and the frequency-spec -> (Wn, btype) mapping against
the reference's cases.
"""
import numpy as np


def filter_mode(freqs):
    """seistorch/signal.py:25-37: number or 1-list -> lowpass, 2-list ->
    bandpass, 'all' -> no filtering."""
    if isinstance(freqs, str):
        if freqs == "all":
            return "all"
        raise ValueError(f"unknown frequency spec {freqs!r}")
    if isinstance(freqs, (int, float)):
        return "lowpass"
    if isinstance(freqs, (list, tuple)):
        if len(freqs) == 1:
            return "lowpass"
        if len(freqs) == 2:
            return "bandpass"
    raise ValueError(f"unsupported frequency spec {freqs!r}")


def butter_ba(order, freqs, dt):
    """(b, a) of seistorch/signal.py:67-75 for a lowpass or bandpass spec."""
    from scipy import signal
    mode = filter_mode(freqs)
    f = [freqs] if isinstance(freqs, (int, float)) else list(freqs)
    wn = [2.0 * fr / (1.0 / dt) for fr in f]
    wn = wn[0] if len(wn) == 1 else wn
    return signal.butter(order, Wn=wn, btype=mode)


def lfilter(b, a, x):
    """Zero-initial-state recursion along axis 0 of x (nt, ...), FP64."""
    b = np.asarray(b, np.float64)
    a = np.asarray(a, np.float64)
    x = np.asarray(x, np.float64)
    n = max(len(b), len(a))
    b = np.pad(b, (0, n - len(b)))
    a = np.pad(a, (0, n - len(a)))
    an = a / a[0]
    # both sums run oldest term first (the order of the HIP kernel; the
    # recursion of a low-Wn Butterworth (b, a) pair is ill-conditioned, so
    # the order shows at ~1e-6 relative against other orders)
    y = np.zeros_like(x)
    z = np.zeros(x.shape[1:])
    for t in range(x.shape[0]):
        w = b[n - 1] * (x[t - n + 1] if t - n + 1 >= 0 else z)
        for k in range(n - 2, -1, -1):
            w = w + b[k] * (x[t - k] if t - k >= 0 else z)
        w = w / a[0]
        d = z
        if n > 1:
            d = an[n - 1] * (y[t - n + 1] if t - n + 1 >= 0 else z)
            for k in range(n - 2, 0, -1):
                d = d + an[k] * (y[t - k] if t - k >= 0 else z)
        y[t] = w - d
    return y


def filtfilt(b, a, x):
    """torchaudio-style filtfilt along axis 0: forward, then backward in time."""
    y = lfilter(b, a, x)
    return lfilter(b, a, y[::-1])[::-1].copy()


def filter_records(records, freqs, dt, order):
    """records (nt, ...) float32 -> filtered float32 (signal.py:91-96)."""
    if filter_mode(freqs) == "all":
        return np.asarray(records, np.float32)
    b, a = butter_ba(order, freqs, dt)
    return filtfilt(b, a, np.asarray(records, np.float64)).astype(np.float32)
