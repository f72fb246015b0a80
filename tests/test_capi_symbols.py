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


def test_ring_particles_reject_bad_radii_on_the_host():
    """pw_vring_particles runs on the host: a zero, negative or non-finite
    smoothing or ring radius is an argument error before anything reaches
    a device (a zero sigma would make every influence entry 0/0)."""
    import numpy as np
    import pytest

    for R, sig in ((1.0, 0.0), (1.0, -0.1), (1.0, float("nan")), (0.0, 0.19), (float("inf"), 0.19)):
        with pytest.raises(capi.PawanError) as e:
            capi.vring_particles(R, 0.1, 3, 49, sig)
        assert e.value.code == capi.PW_ERR_ARG
    pos, radius, volume = capi.vring_particles(1.0, 0.1, 3, 49, 0.1924)
    assert pos.shape == (49 * 49, 3) and np.all(radius == 0.1924) and np.all(volume > 0)


def test_valu_ceiling_rejects_bad_arguments_on_the_host():
    """pw_valu_ceiling checks its form and occupancy before touching a
    device."""
    lib = capi.load()
    r = ctypes.c_double(0.0)
    assert lib.pw_valu_ceiling(4, 32, ctypes.byref(r), None) == capi.PW_ERR_ARG
    assert lib.pw_valu_ceiling(-1, 32, ctypes.byref(r), None) == capi.PW_ERR_ARG
    assert lib.pw_valu_ceiling(0, 6, ctypes.byref(r), None) == capi.PW_ERR_ARG
    assert lib.pw_valu_ceiling(0, 36, ctypes.byref(r), None) == capi.PW_ERR_ARG
    assert lib.pw_valu_ceiling(0, 32, None, None) == capi.PW_ERR_ARG
