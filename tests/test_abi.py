"""CPU checks of the drop-in boundary: the HIP library loads, exports every
entry point declared in include/bastr.h, and its ctypes mirror matches the
header's struct layouts.  No compute call is made (no GPU here)."""
import ctypes as C
import os
import re

import pytest

from bastr_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bastr.h")


def declared_entry_points(path=HEADER):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bastr_[a-z_]+)\s*\(", text)))


def all_headers():
    d = os.path.join(ROOT, "include")
    return [os.path.join(d, f) for f in sorted(os.listdir(d)) if f.endswith(".h")]


def test_header_declares_expected_entry_points():
    names = declared_entry_points()
    expected = sorted("bastr_" + n for n, _, _ in abi.SIGNATURES + abi.ENSEMBLE_SIGNATURES
                      + abi.NOALLOC_SIGNATURES)
    assert names == expected
    closure = declared_entry_points(os.path.join(ROOT, "include", "bastr_closure.h"))
    assert closure == sorted("bastr_" + n for n, _, _ in abi.CLOSURE_SIGNATURES)


def test_hip_library_exports_every_declared_symbol():
    import bastr_amd
    if not os.path.exists(bastr_amd.LIB_PATH):
        pytest.fail("HIP library not built; run __graft_entry__.build()")
    lib = C.CDLL(bastr_amd.LIB_PATH)
    for header in all_headers():
        for name in declared_entry_points(header):
            assert hasattr(lib, name), (header, name)
    bastr_amd.load()


def test_struct_layouts_match_header():
    # offsets fixed by include/bastr.h (natural alignment, no packing pragmas)
    assert C.sizeof(abi.Params) == 72
    assert abi.Params.reynolds.offset == 16
    assert abi.Params.precision.offset == 64
    assert C.sizeof(abi.Diag) == 8 * 18 + 8
    assert abi.Diag.nan_flag.offset == 8 * 18


def test_product_path_has_no_cpu_fallback(monkeypatch, tmp_path):
    import importlib
    import bastr_amd
    monkeypatch.setattr(bastr_amd, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(bastr_amd, "_LIB", None)
    with pytest.raises(RuntimeError):
        bastr_amd.load()
    importlib.reload(bastr_amd)
