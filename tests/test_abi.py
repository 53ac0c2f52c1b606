"""The C-ABI library loads and exports every entry point include/fwhip.h
declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in ("fwhip.h",):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(fwh_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_declared_symbols():
    from adventure_fullwave_amd import fwhip
    lib = ctypes.CDLL(fwhip.LIB_PATH)
    names = declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(fwhip.EXPORTS)


def test_error_without_gpu_is_loud():
    """Without a visible GPU fwh_create fails with an error, never falls back."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from adventure_fullwave_amd import fwhip
    with pytest.raises(fwhip.FwhError):
        fwhip.Context(0)
