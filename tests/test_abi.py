"""CPU checks of the C ABI boundary: the product library loads and exports every function
include/nrcps.h declares (no device calls are made)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "nrcps.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(nrc_\w+)\(", src, re.M)))


def test_header_declares_the_binding_list():
    import nrcps_amd as na

    assert _declared() == sorted(na.EXPORTS)


def test_library_exports_every_declared_symbol():
    import nrcps_amd as na

    if not os.path.exists(na.LIB_PATH):
        pytest.skip("libnrcps_hip.so not built")
    lib = ctypes.CDLL(na.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_library_error_strings_and_arg_checks():
    import nrcps_amd as na

    if not os.path.exists(na.LIB_PATH):
        pytest.skip("libnrcps_hip.so not built")
    lib = na.load_library()
    assert lib.nrc_error_string(0) == b"ok"
    assert lib.nrc_error_string(-5).startswith(b"singular")
    # argument validation happens before any device call
    assert lib.nrc_init(0, None) == -1
    assert lib.nrc_fft_exec(None, None, -1) == -1
    assert lib.nrc_ensemble_run(None, 0, 0, 1, None, None, None) == -1
