"""The C-ABI library loads and exports every entry point include/pawan_gpu.h
declares (no compute: runs without a GPU)."""
import ctypes
import os
import re

from pawan_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "pawan_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pw_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_the_binding_list():
    assert declared_symbols() == sorted(capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_error_strings_without_a_device():
    lib = capi.load()
    assert b"gfx950" in lib.pw_version()
    assert isinstance(lib.pw_last_error(), bytes)


def test_null_context_is_an_argument_error():
    lib = capi.load()
    assert lib.pw_solve(None, capi.PW_FP64) == capi.PW_ERR_ARG
    assert lib.pw_destroy(None) == capi.PW_OK
