"""Parity of the HIP record filter (seis_filtfilt) with the FP64 oracle.

The kernel runs the oracle's recursion in FP64 with the same operation order
(both sums oldest term first, no contraction), so the FP32 outputs are
compared bit for bit.  Against torchaudio, whose summation order is its own,
the agreement is limited by the conditioning of the (b, a) recursion
(tests/test_oracle_signal.py measures it against scipy: ~1e-5 of the peak at
the reference's 1-3 Hz band, dt = 1 ms; 1e-9 for well-conditioned filters).
"""
import numpy as np
import pytest
import torch

from oracle import signal as osg

pytestmark = pytest.mark.gpu


def _records(nt, ntr, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(nt)[:, None]
    x = rng.standard_normal((nt, ntr)) * np.exp(-((t - nt / 3) / (nt / 8)) ** 2)
    return x.astype(np.float32)


@pytest.mark.parametrize("order,freqs,dt,nt,ntr", [
    (3, 5.0, 1e-3, 700, 333),           # lowpass, 4 coefficients, ragged trace count
    (3, [1.0, 3.0], 1e-3, 900, 256),    # the reference's template band, 7 coefficients
    (8, [10.0, 40.0], 4e-3, 300, 65),   # 17 coefficients
    (3, 5.0, 1e-3, 3, 40),              # nt shorter than the filter
    (3, 5.0, 1e-3, 1, 7),               # one sample
])
def test_filtfilt_matches_oracle_bitwise(order, freqs, dt, nt, ntr):
    from seistorch_amd import signal as sg
    b, a = sg.butter(order, freqs, dt)
    x = _records(nt, ntr, nt + ntr)
    ref = osg.filtfilt(b, a, x.astype(np.float64)).astype(np.float32)
    got = sg.filtfilt(torch.from_numpy(x).cuda(), b, a).cpu().numpy()
    diff = np.abs(got - ref).max()
    print(f"filtfilt order={order} freqs={freqs} nt={nt} ntr={ntr} ncoef={len(b)}: "
          f"max |HIP - oracle| {diff:.3e} (peak {np.abs(ref).max():.3e})")
    assert np.array_equal(got, ref)


def test_filtfilt_scaling_only_and_in_place():
    """ncoef = 1 is a pure scaling b0/a0 twice; y may alias x."""
    from seistorch_amd import _lib
    from seistorch_amd import signal as sg
    x = torch.from_numpy(_records(50, 70, 1)).cuda()
    y = sg.filtfilt(x, [3.0], [2.0])
    assert torch.equal(y, (x.double() * 1.5 * 1.5).float())
    b, a = sg.butter(3, [1.0, 3.0], 1e-3)
    ref = sg.filtfilt(x, b, a)
    work = torch.empty(x.shape, dtype=torch.float64, device="cuda")
    import ctypes
    _lib.call("seis_filtfilt", _lib.ptr(x), _lib.ptr(x), _lib.ptr(work), 50, 70,
              b.ctypes.data_as(ctypes.c_void_p), a.ctypes.data_as(ctypes.c_void_p), len(b),
              _lib.stream_of(x))
    assert torch.equal(x, ref)


def test_filtfilt_gradient_is_the_filter_of_the_cotangent():
    """Dot-product test through autograd: <F x, g> == <x, dF^T g>."""
    from seistorch_amd import signal as sg
    b, a = sg.butter(3, [2.0, 6.0], 2e-3)
    x = torch.from_numpy(_records(400, 96, 2)).cuda().requires_grad_(True)
    g = torch.from_numpy(_records(400, 96, 3)).cuda()
    y = sg.filtfilt(x, b, a)
    (y * g).sum().backward()
    lhs = float((y.double() * g.double()).sum())
    rhs = float((x.detach().double() * x.grad.double()).sum())
    print(f"filtfilt dot test: {lhs:.9e} vs {rhs:.9e}")
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + abs(rhs))
    assert torch.equal(x.grad, sg.filtfilt(g, b, a))


def test_seissignal_torch_backend_on_records():
    """SeisSignal.filter(backend='torch') on (nt, R, nch) records and on a
    list of shots, as the FWI loop calls it (seistorch_dist.py:222-225)."""
    from seistorch_amd.signal import SeisSignal
    ss = SeisSignal(dt=1e-3, filter_ord=3)
    rec = _records(600, 2 * 90, 4).reshape(600, 90, 2)
    got = ss.filter(torch.from_numpy(rec).cuda(), [1.0, 3.0], backend="torch").cpu().numpy()
    ref = osg.filter_records(rec, [1.0, 3.0], 1e-3, 3)
    assert np.array_equal(got, ref)
    shots = [torch.from_numpy(rec[:, :45]).cuda().contiguous(),
             torch.from_numpy(rec[:, 45:]).cuda().contiguous()]
    outs = ss.filter(shots, 5.0, backend="torch")
    for o, r in zip(outs, (rec[:, :45], rec[:, 45:])):
        assert np.array_equal(o.cpu().numpy(), osg.filter_records(r, 5.0, 1e-3, 3))
    assert ss.filter(shots, "all", backend="torch") is shots
