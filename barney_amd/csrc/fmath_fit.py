"""Fits the polynomial coefficients of fmath.h (sine and cosine on
[-pi/4, pi/4], arctangent on [0, tan(pi/8)], arcsine on [0, 1/2]) and
prints them.

Method: Cody-Waite reduction by pi/2 (W. J. Cody and W. Waite, Software
Manual for the Elementary Functions, 1980), then
    sin r = r + r^3 P(r^2),   cos r = 1 + r^2 Q(r^2),
with P and Q fitted by weighted least squares on Chebyshev nodes in float64
(weights make the error relative to sin r and cos r), rounded to float32;
likewise atan t = t + t^3 A(t^2) and asin t = t + t^3 B(t^2), A and B of
degree 4.
Run: python barney_amd/csrc/fmath_fit.py
"""
import numpy as np

R = np.pi / 4


def fit(f, deg, weight, n=4000):
    # Chebyshev nodes in u = r^2 over [0, R^2]
    k = np.arange(n)
    u = 0.5 * R * R * (1 - np.cos(np.pi * (k + 0.5) / n))
    V = np.stack([u ** j for j in range(deg + 1)], 1)
    w = weight(u)
    c, *_ = np.linalg.lstsq(V * w[:, None], f(u) * w, rcond=None)
    return c.astype(np.float32)


def sin_p(u):
    r = np.sqrt(u)
    out = np.empty_like(u)
    small = r < 1e-3
    out[small] = -1 / 6 + u[small] / 120 - u[small] ** 2 / 5040
    rr = r[~small]
    out[~small] = (np.sin(rr) - rr) / rr ** 3
    return out


def cos_q(u):
    out = np.empty_like(u)
    small = u < 1e-6
    out[small] = -0.5 + u[small] / 24 - u[small] ** 2 / 720
    uu = u[~small]
    out[~small] = (np.cos(np.sqrt(uu)) - 1) / uu
    return out


def coefficients():
    P = fit(sin_p, 3, lambda u: np.maximum(u, 1e-30) ** 1.5 / np.sin(np.sqrt(u) + 1e-30))
    Q = fit(cos_q, 4, lambda u: u / np.cos(np.sqrt(u)))
    return P, Q


TAN_PI_8 = np.tan(np.pi / 8)


def _odd_tail(f, series):
    """(f(t) - t) / t^3 as a function of u = t^2, by its series near 0."""
    def g(u):
        t = np.sqrt(u)
        out = np.empty_like(u)
        small = t < 1e-3
        out[small] = series(u[small])
        tt = t[~small]
        out[~small] = (f(tt) - tt) / tt ** 3
        return out
    return g


def fit_interval(f, deg, weight, hi2, n=4000):
    k = np.arange(n)
    u = 0.5 * hi2 * (1 - np.cos(np.pi * (k + 0.5) / n))
    V = np.stack([u ** j for j in range(deg + 1)], 1)
    w = weight(u)
    c, *_ = np.linalg.lstsq(V * w[:, None], f(u) * w, rcond=None)
    return c.astype(np.float32)


def inverse_coefficients():
    A = fit_interval(_odd_tail(np.arctan, lambda u: -1 / 3 + u / 5 - u * u / 7), 4,
                     lambda u: np.maximum(u, 1e-30) ** 1.5 / np.arctan(np.sqrt(u) + 1e-30), TAN_PI_8 ** 2)
    B = fit_interval(_odd_tail(np.arcsin, lambda u: 1 / 6 + 3 * u / 40 + 15 * u * u / 336), 4,
                     lambda u: np.maximum(u, 1e-30) ** 1.5 / np.arcsin(np.sqrt(u) + 1e-30), 0.25)
    return A, B


# pi to 60 decimals; float64's own pi is too short for the third part
PI_DIGITS = "3.141592653589793238462643383279502884197169399375105820974944"


def split_half_pi():
    """pi/2 as three float32 parts: hi is the float32 nearest pi/2, mid and
    lo the float32 nearest each remainder, computed in exact decimal
    arithmetic."""
    from decimal import Decimal, getcontext
    getcontext().prec = 80
    half_pi = Decimal(PI_DIGITS) / 2
    hi = np.float32(float(half_pi))
    rem = half_pi - Decimal(float(hi))
    mid = np.float32(float(rem))
    rem = rem - Decimal(float(mid))
    lo = np.float32(float(rem))
    return hi, mid, lo


if __name__ == "__main__":
    P, Q = coefficients()
    hi, mid, lo = split_half_pi()
    A, B = inverse_coefficients()
    for name, c in (("S", P), ("C", Q), ("A", A), ("B", B)):
        for j, v in enumerate(c):
            print(f"BNF_{name}{j + 1} = {float(v).hex()}  // {float(v):.9g}")
    for name, v in (("PIO2_HI", hi), ("PIO2_MID", mid), ("PIO2_LO", lo)):
        print(f"BNF_{name} = {float(v).hex()}  // {float(v):.9g}")
    print(f"BNF_TWO_OVER_PI = {float(np.float32(2 / np.pi)).hex()}")
    print(f"BNF_TAN_PI_8 = {float(np.float32(TAN_PI_8)).hex()}")
