"""The C ABI library loads on a machine without a GPU, exports every entry
point include/tira_hip.h declares, and reports errors cleanly (no compute
calls here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tira_hip.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tira_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("tira_tensorvote3", "tira_evals3_symmetric", "tira_evecs3spherical_symmetric",
              "tira_evals2_symmetric", "tira_evecs2polar_symmetric", "tira_tensorvote3_tk"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from tiralib_amd import _lib
    L = C.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, f"declared but not exported: {missing}"


def test_python_binding_covers_header():
    from tiralib_amd import _lib
    bound = set(_lib.exported_symbols())
    missing = [s for s in declared_symbols() if s not in bound]
    assert not missing, f"no ctypes signature for {missing}"


def test_abi_version_and_error_path_without_compute():
    from tiralib_amd import _lib
    L = _lib.load()
    assert L.tira_hip_abi_version() == 1
    # argument validation happens before any device work
    with pytest.raises(_lib.TiraError):
        _lib.call("tira_tensorvote3", None, None, 4, 4, 4, 1.0, 0.0, 3, 1, 0, 1, 1, 20, 0)
    assert isinstance(L.tira_hip_last_error(), bytes)


def test_python_api_validates_shapes():
    from tiralib_amd import api
    with pytest.raises(ValueError):
        api.tensorvote3(np.zeros((4, 4, 9), np.float32), 1.0)
    with pytest.raises(ValueError):
        api.tensorvote3(np.zeros((4, 4, 4, 9), np.float32), 0.0)
    with pytest.raises(ValueError):
        api.evals3_symmetric(np.zeros(10, np.float32))
