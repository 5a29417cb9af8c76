"""The C-ABI library loads (no GPU needed) and exports every declared symbol."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(seis_\w+)\s*\(", text)
    return names


def test_header_declares_entry_points():
    names = _declared()
    assert "seis_acoustic_forward" in names and "seis_acoustic_backward" in names


def test_library_exports_every_declared_symbol():
    from seistorch_amd import _lib
    lib = _lib.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_declared()) == set(_lib.SIGNATURES), "ctypes table out of sync with the header"
    assert lib.seis_abi_version() == _lib.ABI_VERSION


def test_ring_len_and_error_path():
    from seistorch_amd import _lib
    lib = _lib.load()
    assert lib.seis_ring_len(900, 2500, 49, 0) == 2 * 2400 + 2 * 800
    assert lib.seis_ring_len(100, 100, 49, 0) == -1
    # free surface: no top row, the columns run from row 0 to nz-51
    assert lib.seis_ring_len(850, 2500, 49, 1) == 2400 + 2 * 800
    assert lib.seis_ring_len(51, 101, 49, 1) == 1 + 2 * 1
    assert lib.seis_ring_len(50, 101, 49, 1) == -1
    # argument validation happens on the host, before any device call
    with pytest.raises(RuntimeError, match="batch size"):
        _lib.call("seis_acoustic_step", None, None, None, None, None, 0, 10, 10, 1e-3, 10.0, None)
    with pytest.raises(RuntimeError, match="launch shape"):
        _lib.call("seis_acoustic_forward", None, None, None, 10, None, None, None, 0, None, 0,
                  None, None, None, 0, None, None, None, None, None, None, 1, 120, 120, 1e-3,
                  10.0, 49, 0, 3, 2, 4, None)
    with pytest.raises(RuntimeError, match="null pointer"):
        _lib.call("seis_acoustic_step", None, None, None, None, None, 1, 10, 10, 1e-3, 10.0, None)


def test_batch_and_unet_argument_checks():
    """Host-side validation of the batch planner and the U-Net entry points
    (returns before any device call)."""
    import ctypes
    from seistorch_amd import _lib
    shape = (ctypes.c_int64 * 10)(200, 300, 4, 500, 300, 220, 302, 4, 400, 302)
    dth = (ctypes.c_float * 4)(1e-3, 10.0, 1e-3, 10.0)
    args = [None] * 8 + [0, None, None, None, None, None, 49]  # vp .. snaps, snap_every, u_a .. u_d, table, npml
    with pytest.raises(RuntimeError, match="instance count"):
        _lib.call("seis_acoustic_forward_batch", 0, shape, dth, *args, 4, 2, 4, 0, None)
    with pytest.raises(RuntimeError, match="instance 1: vec_x 4 does not divide nx = 302"):
        _lib.call("seis_acoustic_forward_batch", 2, shape, dth, *args, 4, 2, 4, 0, None)
    with pytest.raises(RuntimeError, match="launch shape"):
        _lib.call("seis_acoustic_forward_batch", 2, shape, dth, *args, 3, 2, 1, 0, None)
    with pytest.raises(RuntimeError, match="null pointer"):
        _lib.call("seis_acoustic_forward_batch", 2, shape, dth, *args, 4, 2, 2, 0, None)
    bad_dt = (ctypes.c_float * 4)(1e-3, 10.0, 0.0, 10.0)
    with pytest.raises(RuntimeError, match="instance 1: dt and h must be positive"):
        _lib.call("seis_acoustic_forward_batch", 2, shape, bad_dt, *args, 4, 2, 2, 0, None)
    # the two-step forward's extra arenas come as a pair (checked before any
    # device call; the addresses are never dereferenced)
    fake = 1 << 20
    no_rec = (ctypes.c_int64 * 10)(200, 300, 4, 500, 0, 220, 302, 4, 400, 0)
    half = [fake] * 6 + [None, None, 0, fake, fake, fake, None, fake, 49]
    with pytest.raises(RuntimeError, match="u_c and u_d are given together"):
        _lib.call("seis_acoustic_forward_batch", 2, no_rec, dth, *half, 4, 2, 2, 0, None)
    with pytest.raises(RuntimeError, match="multiple of 8"):
        _lib.call("seis_unet_conv3x3", None, 1, 16, 16, 12, None, None, 32, 1, None, None, None,
                  None)
    with pytest.raises(RuntimeError, match="even H, W"):
        _lib.call("seis_unet_avgpool2", None, 1, 15, 16, 32, None, None)


def test_filtfilt_argument_checks():
    """Host-side validation of the record filter (before any device call)."""
    import ctypes
    from seistorch_amd import _lib
    b = (ctypes.c_double * 4)(0.1, 0.3, 0.3, 0.1)
    a = (ctypes.c_double * 4)(1.0, -0.5, 0.2, 0.0)
    a0 = (ctypes.c_double * 4)(0.0, -0.5, 0.2, 0.0)
    fake = ctypes.c_void_p(16)  # never dereferenced: validation fails first
    with pytest.raises(RuntimeError, match="null pointer"):
        _lib.call("seis_filtfilt", None, fake, fake, 10, 10, b, a, 4, None)
    with pytest.raises(RuntimeError, match="nt >= 1"):
        _lib.call("seis_filtfilt", fake, fake, fake, 0, 10, b, a, 4, None)
    with pytest.raises(RuntimeError, match="ncoef 33"):
        _lib.call("seis_filtfilt", fake, fake, fake, 10, 10, b, a, 33, None)
    with pytest.raises(RuntimeError, match="a\\[0\\]"):
        _lib.call("seis_filtfilt", fake, fake, fake, 10, 10, b, a0, 4, None)


def test_gradient_postprocess_argument_checks():
    """Host-side validation of the gradient smoothing entry points."""
    import ctypes
    from seistorch_amd import _lib
    w3 = (ctypes.c_float * 3)(0.25, 0.5, 0.25)
    w2 = (ctypes.c_float * 2)(0.5, 0.5)
    f1, f2, f3 = ctypes.c_void_p(16), ctypes.c_void_p(32), ctypes.c_void_p(48)
    with pytest.raises(RuntimeError, match="null pointer"):
        _lib.call("seis_gaussian_filter", None, f2, 10, 10, 0, w3, 3, None)
    with pytest.raises(RuntimeError, match="alias"):
        _lib.call("seis_gaussian_filter", f1, f1, 10, 10, 0, w3, 3, None)
    with pytest.raises(RuntimeError, match="axis"):
        _lib.call("seis_gaussian_filter", f1, f2, 10, 10, 2, w3, 3, None)
    with pytest.raises(RuntimeError, match="smaller than the axis"):
        _lib.call("seis_gaussian_filter", f1, f2, 1, 10, 0, w3, 3, None)
    with pytest.raises(RuntimeError, match="must be odd"):
        _lib.call("seis_smooth_gradient", f1, f2, f3, 10, 10, w2, 2, w3, 3, 1, None, None, None,
                  None)
    with pytest.raises(RuntimeError, match="counts"):
        _lib.call("seis_smooth_gradient", f1, f2, f3, 10, 10, w3, 3, w3, 3, -1, None, None, None,
                  None)
    with pytest.raises(RuntimeError, match="alias"):
        _lib.call("seis_smooth_gradient", f1, f2, f2, 10, 10, w3, 3, w3, 3, 1, None, None, None,
                  None)
    # the no-alloc variant refuses grids beyond its static workspace before any device call
    with pytest.raises(RuntimeError, match="static workspace"):
        _lib.call("seis_smooth_gradient_noalloc", f1, f2, 2049, 2048, w3, 3, w3, 3, 1, None, None,
                  None, None)


def test_variable_density_and_illumination_argument_checks():
    """Host-side validation of the variable-density entry points and of the
    forward's illumination buffer (no device call is reached)."""
    import ctypes
    from seistorch_amd import _lib
    f1, f2 = ctypes.c_void_p(256), ctypes.c_void_p(512)
    with pytest.raises(RuntimeError, match="at least"):
        _lib.call("seis_rho_habc_step", f1, f1, f1, f1, f1, f2, 1, 50, 200, 1e-3, 10.0, 0, 1,
                  None)
    with pytest.raises(RuntimeError, match="null pointer"):
        _lib.call("seis_rho_habc_step", f1, None, f1, f1, f1, f2, 1, 120, 120, 1e-3, 10.0, 0, 1,
                  None)
    with pytest.raises(RuntimeError, match="shots_per_group"):
        _lib.call("seis_rho_habc_forward", f1, f1, f1, f1, 10, f1, f1, 0, None, None, None, 0,
                  f1, None, f1, f1, f1, 1, 120, 120, 1e-3, 10.0, 49, 0, 0, None)
    with pytest.raises(RuntimeError, match="nt must be"):
        _lib.call("seis_rho_habc_forward", f1, f1, f1, f1, 0, f1, f1, 0, None, None, None, 0,
                  f1, None, f1, f1, f1, 1, 120, 120, 1e-3, 10.0, 49, 0, 4, None)
    with pytest.raises(RuntimeError, match="snap_every"):
        _lib.call("seis_rho_habc_forward", f1, f1, f1, f1, 10, f1, f1, 0, None, f1, f2, 0,
                  f1, None, f1, f1, f1, 1, 120, 120, 1e-3, 10.0, 49, 0, 4, None)
    with pytest.raises(RuntimeError, match="null pointer"):
        _lib.call("seis_rho_habc_backward", f1, f1, f1, f1, 10, f1, f1, 0, None, f1, None, 0,
                  f1, f1, f1, f1, f1, f1, f1, None, f1, 1, 120, 120, 1e-3, 10.0, 49, 0, 4, None)
    with pytest.raises(RuntimeError, match="ring storage"):
        _lib.call("seis_rho_habc_backward", f1, f1, f1, f1, 10, f1, f1, 0, None, None, None, 0,
                  f1, f1, f1, f1, f1, f1, f1, f1, f1, 1, 120, 120, 1e-3, 10.0, 49, 0, 4, None)
    with pytest.raises(RuntimeError, match="illum must not alias"):
        _lib.call("seis_acoustic_forward", f1, f1, f1, 10, f1, None, None, 0, None, 0, None, None,
                  None, 0, f1, None, f1, f2, None, None, 1, 120, 120, 1e-3, 10.0, 49, 0, 1, 1, 1,
                  None)
    with pytest.raises(RuntimeError, match="extra-source"):
        _lib.call("seis_acoustic_forward", f1, f1, f1, 10, f1, None, None, 2, None, 0, None, None,
                  None, 0, None, None, f1, f2, None, None, 1, 120, 120, 1e-3, 10.0, 49, 0, 1, 1, 1,
                  None)
    with pytest.raises(RuntimeError, match="one time step per launch"):
        _lib.call("seis_acoustic_forward", f1, f1, f1, 10, f1, f1, f1, 2, None, 0, None, None,
                  None, 0, None, None, f1, f2, ctypes.c_void_p(768), ctypes.c_void_p(1024), 1, 120,
                  120, 1e-3, 10.0, 49, 0, 1, 1, 1, None)
