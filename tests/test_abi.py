"""The C-ABI library loads and exports every symbol include/dcaspt2_amd.h declares (CPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "dcaspt2_amd.h")).read()
    return sorted(set(re.findall(r"\b(dc2_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    from dirac_caspt2_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.EXPORTED)


def test_version_and_error_reporting():
    from dirac_caspt2_amd import _lib
    lib = _lib.lib()
    assert lib.dc2_version().startswith(b"dcaspt2_amd")
    # calls on a null context fail with status 1 and a message, without touching a GPU
    assert lib.dc2_set_spaces(None, 1, 2, 3, 1) == 1
    assert b"null context" in lib.dc2_last_error()


def test_no_device_is_reported():
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK):
        pytest.skip("a GPU is visible")
    from dirac_caspt2_amd import _lib
    h = ctypes.c_void_p()
    assert _lib.lib().dc2_context_create(0, ctypes.byref(h)) != 0


def test_request_footprint_and_server_errors():
    """Host-side parts of the request server: the admission estimate grows with the problem,
    and bad arguments are refused with a message (no GPU needed)."""
    import ctypes
    from dirac_caspt2_amd import _lib
    from dirac_caspt2_amd.api import request_footprint
    small, big = request_footprint(4, 8, 40, 4), request_footprint(16, 12, 164, 6)
    assert 0 < small < big
    lib = _lib.lib()
    h = ctypes.c_void_p()
    assert lib.dc2_server_create(0, 0, 1 << 30, 1, ctypes.byref(h)) == 1
    assert b"invalid argument" in lib.dc2_last_error()
    assert lib.dc2_request_footprint(1, 0, 1, 1, ctypes.byref(ctypes.c_int64())) == 1
