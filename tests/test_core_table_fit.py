"""The branch-free Gaussian core of the tree path's near field: G(s) =
exp(-s/2) B(s) from a table of 64 degree-4 pieces on 0 <= s <= 44.  The
table compiled into pw_pair.hip is read from the source and checked against
a fresh fit, against G by an independent high-precision sum of its series,
and, through the kernel's formulas q' = sqrt(2/pi)(e/3 + s G)/sigma^3 and
f' = sqrt(2/pi) G/sigma^5, against the oracle's KERNEL (gaussian.h:47-60)."""
import math
import os
import re
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "pawan_amd", "csrc")
sys.path.insert(0, CSRC)
import fit_core_table as F  # noqa: E402


def _source():
    return open(os.path.join(CSRC, "pw_pair.hip")).read()


def _compiled_table():
    body = re.search(r"kCoreTable\[64\]\[8\]\s*=\s*\{(.*?)\};", _source(), re.S).group(1)
    rows = re.findall(r"\{([^}]*)\}", body)
    vals = [[float.fromhex(v.strip().rstrip("f")) if "x" in v else float(v.strip().rstrip("f"))
             for v in r.split(",")] for r in rows]
    assert len(vals) == 64 and all(len(r) == 8 for r in vals)
    return np.array(vals, dtype=np.float32)


def test_table_matches_the_fit():
    np.testing.assert_array_equal(_compiled_table(), F.table())
    scale = re.search(r"kCoreScale\s*=\s*([0-9a-fx.p+-]+)f", _source()).group(1)
    assert np.float32(float.fromhex(scale)) == np.float32(F.PIECES / F.S_HI)


def test_g_against_its_series_in_high_precision():
    mp.mp.dps = 40
    rows = _compiled_table()
    s = np.linspace(0.0, 44.0, 1201).astype(np.float32)
    worst = 0.0
    for v in s:
        x = mp.mpf(float(v))
        acc, term, k = mp.mpf(0), mp.mpf(1) / 15, 0
        while term > acc * mp.mpf(10) ** -25 or k < 3:
            acc += term
            term *= x / (2 * k + 7)
            k += 1
        ref = float(mp.e ** (-x / 2) * acc)
        worst = max(worst, abs(float(F.eval32(rows, np.array([v]))[0]) - ref) / ref)
    assert worst < 2.5e-7, worst
    assert F.max_rel_error(rows) < 2.5e-7


def test_pair_factors_against_the_oracle_kernel(oracle):
    """q' = 4 pi q and f' = -(4 pi/3) F of the Gaussian kernel, as the
    kernel forms them in FP32 from the table, e = exp(-s/2) rounded to FP32."""
    rows = _compiled_table()
    sig = 0.7
    k = math.sqrt(2.0 / math.pi)
    for rb in np.linspace(0.02, math.sqrt(44.0), 300):
        rho = rb * sig
        s = np.float32(rb * rb)
        g = float(F.eval32(rows, np.array([s]))[0])
        e = float(np.float32(math.exp(-0.5 * float(s))))
        qp = k * (e / 3.0 + float(s) * g) / sig ** 3
        fp = k * g / sig ** 5
        q, Fk = oracle.kernel(rho, sig)[:2]
        assert abs(qp - 4 * math.pi * q) <= 4e-7 * abs(4 * math.pi * q)
        assert abs(fp + 4 * math.pi / 3 * Fk) <= 4e-7 * abs(4 * math.pi / 3 * Fk)
