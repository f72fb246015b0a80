"""Table of the Gaussian core for the tree path's near-field kernel: the one
function
    G(s) = exp(-s/2) B(s),  B(s) = sum_k s^k / (2k+5)!!,  s = rho^2/sigma^2,
from which that kernel forms both pair factors with no branch,
    4 pi q sigma^3 = sqrt(2/pi) (exp(-s/2)/3 + s G),
    -(4 pi/3) F sigma^5 = sqrt(2/pi) G
(the incomplete-gamma series of erf, A&S 6.5.29, with exp(-s/2) factored:
every term positive).  On 0 <= s <= 44 (the Gaussian core, widened by 1e-4)
G is cut into 64 equal pieces, each a degree-4 polynomial in t = s - s_c
(s_c the piece's centre), a Chebyshev least-squares fit in float64 converted
to the monomial basis and rounded to float32.  A row of the table is
(c0, .., c4, 0, s_c, 0): two 16-byte reads per pair.

G in float64: the series for s < 2 (positive terms, ratio s/(2k+7)); above,
G = (H - e/3)/s with H = exp(-s/2) (1/3 + s B) = s^-3/2 erf(x) / sqrt(2/pi)
- e/s, x = sqrt(s/2), e = exp(-s/2), whose subtraction loses under two
digits.

    python pawan_amd/csrc/fit_core_table.py
"""
import numpy as np
from scipy.special import erf

PIECES = 64
DEG = 4
S_HI = 44.0 * 1.0001
K = np.sqrt(2.0 / np.pi)


def core_g(s) -> np.ndarray:
    s = np.atleast_1d(np.asarray(s, dtype=np.float64))
    out = np.empty_like(s)
    lo = s < 2.0
    sl = s[lo]
    acc = np.zeros_like(sl)
    term = np.full_like(sl, 1.0 / 15.0)
    for k in range(40):
        acc += term
        term = term * sl / (2 * k + 7)
    out[lo] = np.exp(-0.5 * sl) * acc
    sh = s[~lo]
    e = np.exp(-0.5 * sh)
    h = sh ** -1.5 * erf(np.sqrt(0.5 * sh)) / K - e / sh
    out[~lo] = (h - e / 3.0) / sh
    return out


def table(pieces: int = PIECES, deg: int = DEG) -> np.ndarray:
    """(pieces, 8) float32 rows (c0 .. c_deg, s_c, 0 ...)."""
    h = S_HI / pieces
    rows = np.zeros((pieces, 8), dtype=np.float32)
    for k in range(pieces):
        c = np.float32((k + 0.5) * h)
        t = np.cos(np.linspace(0.0, np.pi, 401)) * (h / 2)
        ch = np.polynomial.chebyshev.Chebyshev.fit(t, core_g(float(c) + t), deg, domain=[-h / 2, h / 2])
        p = ch.convert(kind=np.polynomial.Polynomial, domain=[-h / 2, h / 2], window=[-h / 2, h / 2])
        rows[k, :deg + 1] = p.coef.astype(np.float32)
        rows[k, 6] = c
    return rows


def eval32(rows: np.ndarray, s: np.ndarray) -> np.ndarray:
    """G the way the kernel forms it: piece index from s, FP32 Horner in
    t = s - s_c."""
    s = np.asarray(s, dtype=np.float32)
    k = np.minimum((s * np.float32(PIECES / S_HI)).astype(np.int64), PIECES - 1)
    t = (s - rows[k, 6]).astype(np.float32)
    acc = rows[k, DEG].copy()
    for j in range(DEG - 1, -1, -1):
        acc = (acc * t + rows[k, j]).astype(np.float32)
    return acc


def max_rel_error(rows: np.ndarray) -> float:
    s = np.linspace(0.0, 44.0, 400001).astype(np.float32)
    ref = core_g(s.astype(np.float64))
    return float(np.max(np.abs(eval32(rows, s) - ref) / ref))


def hex32(v: float) -> str:
    """A float32 value as a short C hex float literal."""
    if v == 0:
        return "0.0f"
    mant, exp = float(v).hex().split("p")
    mant = mant.rstrip("0")
    if mant.endswith("."):
        mant += "0"
    return f"{mant}p{exp}f"


if __name__ == "__main__":
    rows = table()
    print(f"constexpr float kCoreTable[{PIECES}][8] = {{")
    for r in rows:
        print("    {" + ", ".join(hex32(v) for v in r) + "},")
    print("};")
    print(f"// max relative error (float32 evaluation): {max_rel_error(rows):.3e}")
