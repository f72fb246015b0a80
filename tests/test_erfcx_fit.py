"""The FP32 near-field erf of the pair kernels: erf(x) = 1 - exp(-x^2) R(w),
R the scaled complementary error function as a degree-6 polynomial in
w = sigma/rho.  The coefficients compiled into pw_pair.hip are read from
the source and checked against scipy.special.erfcx over the band they serve
(4 <= rho^2/sigma^2 <= 44), evaluated the way the kernel does (FP32 Horner)."""
import os
import re
import sys

import numpy as np
from scipy.special import erf, erfcx

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "pawan_amd", "csrc")
sys.path.insert(0, CSRC)
import fit_erfcx  # noqa: E402


def _compiled_coefficients():
    src = open(os.path.join(CSRC, "pw_pair.hip")).read()
    body = re.search(r"kErfcx\[7\]\s*=\s*\{([^}]*)\}", src).group(1)
    vals = [float.fromhex(v.strip().rstrip("f")) for v in body.split(",")]
    assert len(vals) == 7
    return np.array(vals, dtype=np.float32)


def test_coefficients_match_the_fit():
    np.testing.assert_array_equal(_compiled_coefficients(), fit_erfcx.coefficients())


def test_scaled_erfc_accuracy():
    assert fit_erfcx.max_rel_error(_compiled_coefficients()) < 5e-7


def test_erf_through_the_kernel_formula():
    """erf = 1 - e R with e = exp(-s/2), s = rho^2/sigma^2 in [4, 44]."""
    cf = _compiled_coefficients()
    s = np.linspace(4.0, 44.0, 20001)
    w = (1.0 / np.sqrt(s)).astype(np.float32)
    e = np.exp(-0.5 * s).astype(np.float32)
    approx = (np.float32(1.0) - e * fit_erfcx.horner32(cf, w)).astype(np.float64)
    exact = erf(np.sqrt(0.5 * s))
    assert np.max(np.abs(approx - exact)) < 1.2e-7   # about 2 FP32 ulps of 1
    # the asymptotic slope: R(w) -> sqrt(2/pi) w as w -> 0
    assert abs(cf[1] - np.sqrt(2.0 / np.pi)) < 0.01
    assert np.all(np.diff(erfcx(1.0 / (np.sqrt(2.0) * np.sort(w.astype(np.float64))))) > 0)
