"""Coefficients of the FP32 near-field erf: erfc(x) = exp(-x^2) R(w) with
R = erfcx = exp(x^2) erfc(x) written as a polynomial in w = 1/(sqrt(2) x),
i.e. w = sigma/rho, on the band the closed form serves (4 <= rho^2/sigma^2
<= 44, widened by 1 % at both ends).  Chebyshev least-squares fit in float64
(scipy.special.erfcx as the function), converted to the monomial basis and
rounded to float32; prints the coefficients (c0 .. c6) as C hex floats and
the maximum relative error of the float32 Horner evaluation.

    python pawan_amd/csrc/fit_erfcx.py
"""
import numpy as np
from scipy.special import erfcx

DEG = 6
W_LO, W_HI = 0.99 / np.sqrt(44.0), 1.01 * 0.5


def coefficients(deg: int = DEG) -> np.ndarray:
    w = np.cos(np.linspace(0.0, np.pi, 8001)) * (W_HI - W_LO) / 2 + (W_HI + W_LO) / 2
    c = np.polynomial.chebyshev.Chebyshev.fit(w, erfcx(1.0 / (np.sqrt(2.0) * w)), deg, domain=[W_LO, W_HI])
    p = c.convert(kind=np.polynomial.Polynomial, domain=[W_LO, W_HI], window=[W_LO, W_HI])
    return p.coef.astype(np.float32)


def horner32(coef: np.ndarray, w: np.ndarray) -> np.ndarray:
    w = w.astype(np.float32)
    acc = np.full_like(w, coef[-1])
    for k in range(len(coef) - 2, -1, -1):
        acc = (acc * w + coef[k]).astype(np.float32)
    return acc


def max_rel_error(coef: np.ndarray) -> float:
    w = np.linspace(W_LO, W_HI, 400001)
    ref = erfcx(1.0 / (np.sqrt(2.0) * w.astype(np.float32).astype(np.float64)))
    return float(np.max(np.abs(horner32(coef, w) - ref) / ref))


if __name__ == "__main__":
    cf = coefficients()
    for k, v in enumerate(cf):
        print(f"c{k} = {float(v).hex()}  ({float(v):.9g})")
    print(f"max relative error (float32 Horner): {max_rel_error(cf):.3e}")
