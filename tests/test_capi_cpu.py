"""CPU-side checks of the C-ABI library: it loads and exports every declared symbol."""
import ctypes
import os
import subprocess

import pytest

from timeevolver_amd import capi


def test_library_exports_header_symbols():
    L = capi.load()
    syms = capi.header_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(capi._SIGS) <= set(syms)


def test_version_and_no_device():
    L = capi.load()
    assert L.te_version() == 201                     # core/version.h:7
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    assert L.te_ctx_create(0, ctypes.byref(h)) != capi.TE_OK


def test_header_is_plain_c():
    """The boundary header compiles as C (no C++ types cross the ABI)."""
    r = subprocess.run(["gcc", "-x", "c", "-std=c99", "-fsyntax-only", "-Wall", "-Werror", capi.HEADER],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
