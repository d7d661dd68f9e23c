"""C-ABI library: it loads and exports every function include/dwba_hip.h declares (no GPU needed)."""
import ctypes
import os
import re

from dwba_amd import _hip
from dwba_amd._build import lib_path

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "dwba_hip.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(dwba_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_library_exists():
    assert os.path.exists(lib_path()), "run __graft_entry__.build() first"


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(lib_path())
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_functions()) == set(_hip.exported_symbols())


def test_version_and_errors_without_gpu():
    lib = _hip.load_library()
    assert lib.dwba_version() >= 100
    # shape queries are pure host functions
    assert lib.dwba_il_plain_f64_workspace_bytes(100, 100, 30) > 0
    assert lib.dwba_il_plain_f64_workspace_bytes(101, 100, 30) == 0   # n_rows > n_cols
    assert lib.dwba_il_plain_f64_workspace_bytes(10, 10, 64) == 0     # L_max too large
