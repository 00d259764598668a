"""Least-squares fits behind the polynomial coefficients of
include/aarbf_detmath.h (the renderer's deterministic logarithm and
exponential), in 60-digit decimal arithmetic:

  log:  ln(1 + f) = f + f^2 q(f),  q of degree 8, fitted to
        (ln(1 + f) - f) / f^2 on 64 Chebyshev nodes of
        [sqrt(2)/2 - 1, sqrt(2) - 1] (the reduced mantissa range);
  exp:  e^r = 1 + r + r^2 P(r),  P of degree 5, fitted to
        (e^r - 1 - r) / r^2 on 64 Chebyshev nodes of [-ln2/2, ln2/2].

Each fit solves the normal equations exactly in decimal arithmetic; the
coefficients are then rounded to binary32.  Run it to print them:

    python tests/golden/fit_detmath.py

tests/test_oracle.py checks that the header holds exactly these values.
"""
from decimal import Decimal as D, getcontext

import numpy as np

getcontext().prec = 60
NODES = 64


def _arctan_inv(n):
    x = D(1) / n
    s, t, k, sign = D(0), x, 1, 1
    while abs(t) > D(10) ** -70:
        s += sign * t / k
        t, k, sign = t / (n * n), k + 2, -sign
    return s


PI = 4 * (4 * _arctan_inv(5) - _arctan_inv(239))  # Machin's formula


def _cos(x):
    s, t, n = D(0), D(1), 0
    while abs(t) > D(10) ** -70:
        s += t
        n += 2
        t = -t * x * x / (n * (n - 1))
    return s


def _exp(x):
    s, t, n = D(0), D(1), 0
    while abs(t) > D(10) ** -70:
        s += t
        n += 1
        t = t * x / n
    return s


def _lstsq(A, y):
    """Least squares through the normal equations, Gaussian elimination
    with partial pivoting, all in decimal arithmetic."""
    n = len(A[0])
    M = [[sum(A[k][i] * A[k][j] for k in range(len(A))) for j in range(n)] for i in range(n)]
    v = [sum(A[k][i] * y[k] for k in range(len(A))) for i in range(n)]
    for i in range(n):
        p = max(range(i, n), key=lambda r: abs(M[r][i]))
        M[i], M[p], v[i], v[p] = M[p], M[i], v[p], v[i]
        for r in range(i + 1, n):
            f = M[r][i] / M[i][i]
            for c in range(i, n):
                M[r][c] -= f * M[i][c]
            v[r] -= f * v[i]
    x = [D(0)] * n
    for i in range(n - 1, -1, -1):
        x[i] = (v[i] - sum(M[i][c] * x[c] for c in range(i + 1, n))) / M[i][i]
    return x


def _chebyshev(a, b, n):
    return [(a + b) / 2 + (b - a) / 2 * _cos((2 * k + 1) * PI / (2 * n)) for k in range(n)]


def log_coefficients():
    """q_0 .. q_8 (q_0 is the constant term) as binary32."""
    fs = _chebyshev(D(2).sqrt() / 2 - 1, D(2).sqrt() - 1, NODES)
    ys = [((1 + f).ln() - f) / (f * f) for f in fs]
    return np.array([float(c) for c in _lstsq([[f ** j for j in range(9)] for f in fs], ys)], np.float32)


def exp_coefficients():
    """P_0 .. P_5 (P_0 is the constant term) as binary32."""
    ln2 = D(2).ln()
    rs = _chebyshev(-ln2 / 2, ln2 / 2, NODES)
    ys = [(_exp(r) - 1 - r) / (r * r) for r in rs]
    return np.array([float(c) for c in _lstsq([[r ** j for j in range(6)] for r in rs], ys)], np.float32)


def c_literal(v) -> str:
    """The binary32 value as the header writes it: its exact double value,
    shortest round-trip form, with an f suffix."""
    return repr(float(np.float32(v))) + "f"


if __name__ == "__main__":
    print("log q (highest degree first):", ", ".join(c_literal(v) for v in log_coefficients()[::-1]))
    print("exp P (highest degree first):", ", ".join(c_literal(v) for v in exp_coefficients()[::-1]))
