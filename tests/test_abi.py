"""The C-ABI library loads without a GPU and exports every entry point that include/subpx_hip.h
declares, each with a ctypes signature in imageprocessor_new_amd._lib (no compute calls)."""
import ctypes
import re
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "subpx_hip.h"


def _declared():
    text = re.sub(r"/\*.*?\*/", " ", HEADER.read_text(), flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(spx_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def _library():
    from imageprocessor_new_amd import _lib

    if not _lib.LIB_PATH.exists():
        subprocess.run([sys.executable, "-m", "imageprocessor_new_amd._build"], cwd=ROOT, check=True)
    return _lib, ctypes.CDLL(str(_lib.LIB_PATH))


def test_header_declares_the_entry_points():
    names = _declared()
    assert len(names) >= 30
    for must in ("spx_fit_batch", "spx_threshold_cc", "spx_hole_sweep", "spx_outline_windows", "spx_cnn_propose"):
        assert must in names


def test_every_declared_symbol_is_exported_and_typed():
    _lib, lib = _library()
    names = _declared()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    untyped = [n for n in names if n not in _lib.SIGNATURES]
    assert not untyped, f"declared but without a ctypes signature: {untyped}"
    extra = [n for n in _lib.SIGNATURES if n not in names]
    assert not extra, f"typed but not declared in the header: {extra}"


def test_load_declares_signatures_and_reports_version():
    _lib, _ = _library()
    lib = _lib.load()
    assert lib.spx_fit_batch.restype is ctypes.c_int32
    v = lib.spx_version()
    assert v is not None
