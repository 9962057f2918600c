"""CPU checks of the drop-in boundary: libshellgpu.so loads without a GPU and
exports every entry point that include/shell_gpu.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "shell_gpu.h")
LIB = os.path.join(ROOT, "shell-encryption_amd", "libshellgpu.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(shg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("shg_create", "shg_ntt_fwd", "shg_ntt_inv", "shg_relin_apply",
              "shg_galois_apply", "shg_mod_reduce_lsb", "shg_substitute", "shg_tensor_product"):
        assert s in syms


@pytest.mark.skipif(not os.path.exists(LIB), reason="libshellgpu.so not built")
def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libshellgpu.so not built")
def test_error_paths_without_device():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "shell-encryption_amd", "python"))
    import shell_gpu as sg
    L = sg.lib()
    # Null context and null arguments are rejected before any device call.
    assert L.shg_sync(None) == -1
    assert L.shg_ntt_fwd(None, None, 1, 1, None, 0) == -1
    assert b"context" in L.shg_last_error()
    # Entry points added with the lazy absorbs, prepared keys and the autotuner.
    assert L.shg_fused_mul_sum_add(None, 0, None, None, None, None, 1, 1, None) == -1
    assert L.shg_autotune_ntt_moduli(None, None, 1, 1, None) == -1
    assert L.shg_ks_key_create(None, None, None, 1, 1, None, None) == -1
    assert L.shg_ks_key_destroy(None, None) == -1
    assert L.shg_relin_apply_key(None, None, 2, 7, None, None, None, None, 1, 1, None) == -1
    assert b"key" in L.shg_last_error()
    assert L.shg_galois_apply_key(None, None, 1, None, None, None, None, 1, 1, None) == -1
