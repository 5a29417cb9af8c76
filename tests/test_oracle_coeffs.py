"""Closed-form values for the PML/HABC coefficients and the Ricker wavelet."""
import numpy as np

from oracle import coeffs


def test_pml_closed_form():
    prof = coeffs.pml_profile(50)
    assert abs(prof[0] - 180.0) < 1e-12          # d0 = 1.5*1500/50*log10(1e4)
    assert abs(prof[49] - 180.0 / 2500.0) < 1e-12  # 0.072
    assert prof[50] == 0.0
    d = coeffs.pml_coefficients_2d(300, 500, 50)
    assert abs(d[0, 0] - np.sqrt(2) * 180.0) < 1e-9
    assert abs(d[-1, -1] - np.sqrt(2) * 180.0) < 1e-9
    assert np.all(d[50:-50, 50:-50] == 0.0)
    assert abs(d[150, 0] - 180.0) < 1e-12 and abs(d[0, 250] - 180.0) < 1e-12
    np.testing.assert_allclose(d[49, 200], 0.072, rtol=1e-12)
    np.testing.assert_allclose(d[:, ::-1], d, rtol=0, atol=0)  # mirror symmetric
    np.testing.assert_allclose(d[::-1, :], d, rtol=0, atol=0)


def test_pml_multiple_free_surface():
    d = coeffs.pml_coefficients_2d(200, 300, 50, multiple=True)
    assert np.all(d[0, 60:-60] == 0.0)      # no top strip
    assert abs(d[-1, 150] - 180.0) < 1e-12  # bottom strip present


def test_habc_masks_and_weights():
    top, bottom, left, right = coeffs.habc_masks(120, 140, 50)
    for r in range(50):
        cols = np.nonzero(top[r])[0]
        assert cols[0] == r and cols[-1] == 139 - r
    np.testing.assert_array_equal(bottom, top[::-1])
    np.testing.assert_array_equal(right, left[:, ::-1])
    for j in range(50):
        rows = np.nonzero(left[:, j])[0]
        assert rows[0] == j and rows[-1] == 119 - j
    b = coeffs.habc_weights(50)
    np.testing.assert_allclose(b, (49 - np.arange(50)) / 49.0, rtol=1e-13)
    d = coeffs.habc_coefficients_2d(120, 140)
    assert d[0, 70] == 1.0 and d[49, 70] == 0.0 and d[70, 0] == 1.0
    assert np.all(d[50:-50, 50:-50] == 0)


def test_ricker_closed_form():
    fm, dt, delay = 10.0, 1e-3, 100
    w = coeffs.ricker(fm, dt, 400, delay=delay)
    assert w[delay] == 1.0 and np.argmax(w) == delay
    # zero crossings at |t| = 1/(sqrt(2) pi fm)
    t0 = 1 / (np.sqrt(2) * np.pi * fm)
    k = int(np.floor(t0 / dt))
    assert w[delay + k] > 0 and w[delay + k + 1] < 0
    np.testing.assert_allclose(w[delay - 50:delay], w[delay + 50:delay:-1], rtol=1e-12)
    # minima at t = +-sqrt(3/2)/(pi fm) with value -2 exp(-3/2)
    assert abs(w.min() + 2 * np.exp(-1.5)) < 2e-3


# ------------------------------------------------------- the product's generators

def test_product_pml_matches_oracle():
    """seistorch_amd.pml (FP32, the reference's op order) against the FP64
    restatement pinned above, with and without a free surface."""
    import torch
    from seistorch_amd.pml import generate_pml_coefficients_2d
    for mult in (False, True):
        for nz, nx in ((300, 500), (151, 207)):
            d = generate_pml_coefficients_2d((nz, nx), 50, multiple=mult)
            assert d.dtype == torch.float32 and d.shape == (nz, nx)
            ref = coeffs.pml_coefficients_2d(nz, nx, 50, multiple=mult)
            np.testing.assert_allclose(d.numpy(), ref, rtol=2e-7, atol=1e-5)
            assert np.all(d.numpy()[ref == 0.0] == 0.0)
    d = generate_pml_coefficients_2d((300, 500), 50).numpy()
    assert abs(d[0, 250] - 180.0) < 2e-5 and abs(d[49, 250] - 0.072) < 1e-7


def test_product_habc_matches_oracle():
    from seistorch_amd.habc import generate_habc_coefficients_2d
    for mult in (False, True):
        b = generate_habc_coefficients_2d((130, 170), 50, multiple=mult).numpy()
        ref = coeffs.habc_coefficients_2d(130, 170, multiple=mult)
        np.testing.assert_allclose(b, ref, rtol=0, atol=1e-7)


def test_product_ricker_matches_oracle():
    from seistorch_amd.utils import ricker_wave
    for fm, dt, nt, delay in ((10.0, 1e-3, 1500, 256), (8.0, 8e-4, 5000, 200),
                              (25.0, 5e-4, 300, 40)):
        w = ricker_wave(fm, dt, nt, delay=delay).numpy()
        ref = coeffs.ricker(fm, dt, nt, delay=delay)
        assert w.dtype == np.float32
        np.testing.assert_allclose(w, ref, rtol=0, atol=6e-8)
        assert w[delay] == 1.0
    wi = ricker_wave(10.0, 1e-3, 300, delay=80, inverse=True).numpy()
    np.testing.assert_allclose(wi, coeffs.ricker(10.0, 1e-3, 300, delay=80, inverse=True),
                               rtol=0, atol=6e-8)
