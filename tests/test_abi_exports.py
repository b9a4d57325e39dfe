"""The C-ABI library loads and exports every entry point declared in include/spike_hip.h
and include/spike_mlp.h (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

from spike_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for h in ("spike_hip.h", "spike_mlp.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        syms |= set(re.findall(r"\b(spk_[a-z0-9_]+)\s*\(", text))
    return sorted(syms)


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "spk_run" in syms and "spk_create" in syms and "spk_mlp_train" in syms
    assert len(syms) >= 40


def test_library_exports_all_declared_symbols():
    assert os.path.exists(abi.LIB_PATH), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(abi.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"missing exports: {missing}"


def test_python_binding_covers_header():
    assert set(declared_symbols()) <= set(abi.SIGNATURES), \
        sorted(set(declared_symbols()) - set(abi.SIGNATURES))
