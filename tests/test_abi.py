"""The C-ABI library loads without a GPU and exports every entry point include/*.h declares;
the ctypes layer binds each of them."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:int|const char\*|void)\s+(petch_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "petch_ctx_create" in names and len(names) >= 8


def test_library_exports_every_declared_symbol():
    from petch import _abi
    if not os.path.exists(_abi.LIB_PATH):
        pytest.fail("libpetch_hip.so is not built (python plasma-etching-code_amd/build.py)")
    lib = ctypes.CDLL(_abi.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_ctypes_layer_binds_every_declared_symbol():
    from petch import _abi
    bound = set(_abi.SIGNATURES) | {"petch_last_error"}
    assert set(declared_symbols()) <= bound, sorted(set(declared_symbols()) - bound)
    L = _abi.lib()
    assert L.petch_version() >= 1


def test_ctx_create_rejects_bad_device_without_gpu():
    from petch import _abi
    h = ctypes.c_void_p()
    assert _abi.lib().petch_ctx_create(-1, ctypes.byref(h)) == _abi.PETCH_E_ARG
