#!/usr/bin/env python3
"""Fits the FP32 polynomial coefficients of the walk kernel's math
(pcs_device.h; the oracle carries the same constants) and prints them as
hex floats.

- ln(1+f) = f + f^2 P(f), P of degree 6, on [sqrt(1/2)-1, sqrt(2)-1]
  (weighted least squares on Chebyshev nodes, reweighted iteratively by the
  residual - Lawson's algorithm - to approach the minimax polynomial, then
  rounded to FP32);
- 2^f = 1 + f Q(f), Q of degree 5, on [0, 1) (same method, weighted for the
  relative error of 2^f; the constant 1 keeps 2^0 exact);
- the two Box-Muller angle tables for the angle 2 pi (a + 1/2) / 2^16 of a
  16-bit code a = i << 8 | r, which is the sector centre 2 pi (i + 1/2) / 256
  plus beta = 2 pi (r - 127.5) / 2^16: SINCOS_TAB holds sin and cos of the
  sector centres (i = 0..255), BETA_TAB holds sin beta and cos beta - 1
  (r = 0..255); every entry computed to 50 digits and rounded to
  nearest-even FP32;
- the FP64 table of 2^(j/64), j = 0..63, for the pmc weights' exponential,
  and its Cody-Waite split of ln2 / 64 (a head of 35 significant bits, so
  that k * head is exact for |k| < 2^18, and the rounded tail), both from 50
  digits, rounded to nearest FP64.
"""
import mpmath
import numpy as np


def lawson(A, y, w0, iters):
    w = w0.copy()
    c, *_ = np.linalg.lstsq(A * w[:, None], y * w, rcond=None)
    for _ in range(iters):
        r = (A @ c - y) * w0
        w = w * (np.abs(r) / np.abs(r).max() + 1e-4) ** 0.6
        w /= w.max()
        c, *_ = np.linalg.lstsq(A * w[:, None], y * w, rcond=None)
    return c


def cheb(a, b, n=6000):
    return 0.5 * (a + b) + 0.5 * (b - a) * np.cos(np.pi * (np.arange(n) + 0.5) / n)


def hexf(v):
    return float.hex(float(np.float32(v))).replace("0000000p", "p")


def f32_rn(x):
    """Nearest-even FP32 value of the mpmath number x, as a Python float."""
    if x == 0:
        return 0.0
    sign = -1.0 if x < 0 else 1.0
    ax = abs(x)
    e = int(mpmath.floor(mpmath.log(ax, 2)))
    if mpmath.mpf(2) ** e > ax:  # guard the floor of the logarithm
        e -= 1
    elif mpmath.mpf(2) ** (e + 1) <= ax:
        e += 1
    m = ax * mpmath.mpf(2) ** (23 - e)  # in [2^23, 2^24)
    fl = int(mpmath.floor(m))
    rem = m - fl
    if rem > 0.5 or (rem == 0.5 and fl % 2 == 1):
        fl += 1
    return sign * float(fl) * 2.0 ** (e - 23)


def sincos_table():
    mpmath.mp.dps = 50
    tab = []
    for i in range(256):
        th = 2 * mpmath.pi * (i + mpmath.mpf("0.5")) / 256
        tab += [f32_rn(mpmath.sin(th)), f32_rn(mpmath.cos(th))]
    return tab


def beta_table():
    mpmath.mp.dps = 50
    tab = []
    for r in range(256):
        b = 2 * mpmath.pi * (r - mpmath.mpf("127.5")) / 65536
        tab += [f32_rn(mpmath.sin(b)), f32_rn(mpmath.cos(b) - 1)]
    return tab


def exp2_64_table():
    mpmath.mp.dps = 50
    return [float(mpmath.power(2, mpmath.mpf(j) / 64)) for j in range(64)]


def ln2_64_split():
    mpmath.mp.dps = 50
    c = mpmath.log(2) / 64
    e = int(mpmath.floor(mpmath.log(c, 2)))
    hi = mpmath.nint(c / mpmath.power(2, e - 34)) * mpmath.power(2, e - 34)  # 35 significant bits
    return float(hi), float(c - hi), float(64 / mpmath.log(2))


def main():
    a, b = np.sqrt(0.5) - 1, np.sqrt(2) - 1
    x = cheb(a, b)
    A = np.stack([x ** p for p in range(7)], 1)
    logp = lawson(A, (np.log1p(x) - x) / x ** 2, np.abs(x) + 1e-12, 80).astype(np.float32)
    print("LOG_P", [hexf(v) for v in logp])
    x = cheb(0.0, 1.0)
    A = np.stack([x ** p for p in range(6)], 1)
    exq = lawson(A, np.expm1(x * np.log(2.0)) / x, x / np.exp2(x), 80).astype(np.float32)
    print("EXP2_Q", [hexf(v) for v in exq])
    print("SINCOS_TAB", [float.hex(v) for v in sincos_table()])
    print("BETA_TAB", [float.hex(v) for v in beta_table()])
    print("EXP2_64_TAB", [float.hex(v) for v in exp2_64_table()])
    print("LN2_64_HI_LO_INV", [float.hex(v) for v in ln2_64_split()])


if __name__ == "__main__":
    main()
