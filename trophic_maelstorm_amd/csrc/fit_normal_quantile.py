"""Fits the coefficients of the FP32 normal quantile in tm_device.h (norm_icdf_f).

    python trophic_maelstorm_amd/csrc/fit_normal_quantile.py

Form (M. Giles, "Approximating the erfinv function", GPU Computing Gems Jade
Edition, 2011, scaled by sqrt(2)): with t = 2p - 1 and w = -ln(4p(1 - p)),

    Phi^{-1}(p) = t * A(w - 5/2)        w < 5
                = t * B(sqrt(w) - 3)    w >= 5

A and B are degree-9 polynomials fitted here by weighted least squares on
Chebyshev nodes of each interval, against scipy's ndtri, reweighting toward the
largest relative errors.  The tail interval ends at p = 2^-25, below the smallest
p that Stream::u32() yields (2^-24).  Prints the FP32 literals (highest degree
first) and the relative errors of the polynomials in FP64.
"""
import numpy as np
from scipy.special import ndtri

W0 = 5.0           # split between the two forms
DEG = 9


def ratio(w):
    """Phi^{-1}(p) / t on the lower root p(w) = (1 - sqrt(1 - e^-w)) / 2."""
    t = -np.sqrt(-np.expm1(-w))
    p = (1.0 + t) / 2.0
    return ndtri(p) / t


def nodes(a, b, n):
    k = np.arange(n)
    return a + (np.cos(np.pi * (k + 0.5) / n) + 1.0) / 2.0 * (b - a)


def fit(x, y, deg, iters=30):
    wts = np.ones_like(x)
    for _ in range(iters):
        V = np.vander(x, deg + 1)
        c, *_ = np.linalg.lstsq(V * (wts / y)[:, None], wts, rcond=None)
        err = np.abs(np.polyval(c, x) / y - 1.0)
        wts = wts * (1.0 + 5.0 * err / err.max())
    return c, err.max()


def main():
    wc = nodes(1e-6, W0, 6000)            # w -> 0 is t -> 0, where ratio -> sqrt(pi/2)
    cc, ec = fit(wc - 2.5, ratio(wc), DEG)
    pmin = 2.0 ** -25
    wmax = -np.log(4.0 * pmin * (1.0 - pmin))
    s = nodes(np.sqrt(W0), np.sqrt(wmax), 6000)
    ct, et = fit(s - 3.0, ratio(s * s), DEG)
    lit = lambda c: ", ".join("%.9ef" % float(np.float32(v)) for v in c)
    print(f"A (w < {W0}), max relative error {ec:.2e}:\n  {lit(cc)}")
    print(f"B (w >= {W0}), max relative error {et:.2e}:\n  {lit(ct)}")


if __name__ == "__main__":
    main()
