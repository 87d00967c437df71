"""The C-ABI library loads and exports every entry point declared in include/svimx.h."""

import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "svimx.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z0-9_]+\s*\*?\s*(svimx_[a-z0-9_]+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "svimx_fractions" in syms and "svimx_nr_exec" in syms and len(syms) >= 20


def test_library_exports_every_declared_symbol():
    from svierosion_amd import _abi
    lib = ctypes.CDLL(str(_abi.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    bound = set(_abi.exported_symbols())
    assert not [s for s in declared_symbols() if s not in bound]


def test_error_mapping_without_gpu():
    from svierosion_amd import _abi
    assert _abi.lib.svimx_version() >= 1
    assert _abi.lib.svimx_workspace_bytes(90, 90, 90) > 4096
    # a null workspace is rejected on the host before any device call
    rc = _abi.lib.svimx_fractions(None, None, None, 0, 10, 10, 10, 0.1, None, None, 5.5, None, 0)
    assert rc == _abi.EBADARG
    assert "workspace" in _abi.lib.svimx_last_error().decode()
