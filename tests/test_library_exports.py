"""The C-ABI library loads and exports every entry point include/xtreme_hip.h declares (no GPU calls)."""
import ctypes
import os
import re

from extreme_amd import xtreme

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "xtreme_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xtr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("xtr_op_create", "xtr_op_build_green", "xtr_op_set_green", "xtr_op_prepare", "xtr_op_apply",
              "xtr_fgmres_solve", "xtr_fgmres_solve_batch_device", "xtr_op_destroy", "xtr_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(xtreme.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_last_error_is_empty_without_calls():
    lib = ctypes.CDLL(xtreme.LIB_PATH)
    buf = ctypes.create_string_buffer(64)
    assert lib.xtr_last_error(buf, 64) == 0
