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


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of the C structs have the header's size and field offsets."""
    import shutil
    import subprocess
    from dwba_amd import batch, xblock
    gcc = shutil.which("gcc")
    if gcc is None:
        import pytest
        pytest.skip("gcc not available")
    structs = {
        "dwba_plain_job": _hip.PlainJob, "dwba_prep_job": _hip.PrepJob, "dwba_tail_job": _hip.TailJob,
        "dwba_xgemm_energy": batch.XgemmEnergy, "dwba_dgemv_energy": batch.DgemvEnergy,
        "dwba_ddot_pair": batch.DdotPair, "dwba_finalize_pair": batch.FinalizePair,
        "dwba_lane_grid": batch.LaneGrid, "dwba_filon_group": batch.FilonGroup,
        "dwba_amp_pair": _hip.AmpPair, "dwba_cw_wave": _hip.CwWave, "dwba_cw_finish": _hip.CwFinish,
        "dwba_cw_bypass": _hip.CwBypass, "dwba_scan_energy": _hip.ScanEnergy, "dwba_scan_wave": _hip.ScanWave,
        "dwba_scan_pair": _hip.ScanPair, "dwba_scan_options": _hip.ScanOptions,
        "dwba_scan_pair_plan": _hip.ScanPairPlan, "dwba_xblock_grid": xblock.XBlockGrid,
        "dwba_xblock_group": xblock.XBlockGroup,
    }
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, cls in structs.items():
        src.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            src.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    src.append("return 0;}")
    c = tmp_path / "probe.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "probe"
    subprocess.run([gcc, str(c), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line.strip():
            a, b, v = line.split()
            got[(a, b)] = int(v)
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
