"""The C-ABI library loads and exports every entry point include/llr_hip.h declares (no GPU calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "llr_hip.h")
LIB = os.path.join(ROOT, "llr-reco_amd", "llr_hip", "libllr_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(llr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "llr_lift_data_set" in names and "llr_recon_run" in names
    assert len(names) >= 35


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build libllr_hip.so first (__graft_entry__.build())"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_lists_the_same_symbols():
    import llr_hip._lib as L

    assert sorted(L.EXPORTS) == declared()


def test_library_carries_gfx950_code_object():
    # the embedded offload bundle names its target triple
    assert b"amdgcn-amd-amdhsa--gfx950" in open(LIB, "rb").read()


def test_abi_version_and_error_string_without_gpu():
    import llr_hip._lib as L

    lib = ctypes.CDLL(LIB)
    assert lib.llr_abi_version() == L.ABI_VERSION == 108
    lib.llr_last_error.restype = ctypes.c_char_p
    assert lib.llr_last_error() == b""
    g = (ctypes.c_size_t * 3)()
    lib.llr_grid_dim((ctypes.c_size_t * 3)(256, 256, 128), (ctypes.c_size_t * 3)(8, 8, 8), g)
    assert list(g) == [32, 32, 16]
