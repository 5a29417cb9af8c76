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
    assert lib.seis_ring_len(900, 2500, 49) == 2 * 2400 + 2 * 800
    assert lib.seis_ring_len(100, 100, 49) == -1
    # argument validation happens on the host, before any device call
    with pytest.raises(RuntimeError, match="batch size"):
        _lib.call("seis_acoustic_step", None, None, None, None, None, 0, 10, 10, 1e-3, 10.0, None)
    with pytest.raises(RuntimeError, match="launch shape"):
        _lib.call("seis_acoustic_forward", None, None, None, 10, None, None, 0, None, None, None,
                  0, None, None, 1, 120, 120, 1e-3, 10.0, 49, 3, 2, 4, None)
    with pytest.raises(RuntimeError, match="null pointer"):
        _lib.call("seis_acoustic_step", None, None, None, None, None, 1, 10, 10, 1e-3, 10.0, None)
