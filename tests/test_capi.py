"""The C-ABI library loads and exports every symbol include/cyclix_msparc.h
declares (no device calls: runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def header_symbols():
    text = (ROOT / "include" / "cyclix_msparc.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(cyx_[a-z_0-9]+)\s*\(", text, re.M)))


def test_header_lists_the_python_binding_exports():
    from cyclix_msparc_amd import _native as nat
    assert header_symbols() == sorted(nat.EXPORTS)


def test_library_exports_every_declared_symbol():
    from cyclix_msparc_amd import _native as nat
    if not nat.LIB_PATH.exists():
        pytest.fail(f"{nat.LIB_PATH} not built (run python -m cyclix_msparc_amd.build)")
    lib = ctypes.CDLL(str(nat.LIB_PATH))
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_null_arguments_are_rejected_without_touching_the_device():
    from cyclix_msparc_amd import _native as nat
    lib = nat.load()
    h = ctypes.c_void_p()
    assert lib.cyx_ham_create(None, None, None, None, None, None, ctypes.byref(h)) == nat.CYX_ERR_ARG
    assert b"null" in lib.cyx_last_error()
    assert lib.cyx_h_apply(None, 0, None, None, 1, 1) == nat.CYX_ERR_ARG
    assert lib.cyx_poisson_create(None, None, ctypes.byref(h)) == nat.CYX_ERR_ARG
    assert lib.cyx_version().startswith(b"cyclix_msparc_amd")
