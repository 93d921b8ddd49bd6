#!/usr/bin/env python3
"""Fits the FP32 polynomial coefficients of the walk kernel's math
(pcs_device.h; the oracle carries the same constants) and prints them as
hex floats.

- ln(1+f) = f + f^2 P(f), P of degree 7, on [sqrt(1/2)-1, sqrt(2)-1];
- sin(2 pi t) = t S(t^2), S of degree 3, and cos(2 pi t) = 1 + t^2 C(t^2),
  C of degree 3, on |t| <= 1/8 turn.

Method: weighted least squares on Chebyshev nodes, reweighted iteratively
by the residual (Lawson's algorithm) to approach the minimax polynomial,
then rounded to FP32.
"""
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


def main():
    a, b = np.sqrt(0.5) - 1, np.sqrt(2) - 1
    x = cheb(a, b)
    A = np.stack([x ** p for p in range(8)], 1)
    logp = lawson(A, (np.log1p(x) - x) / x ** 2, np.abs(x) + 1e-12, 80).astype(np.float32)
    t = cheb(0, 0.125)
    t2 = t * t
    A = np.stack([t2 ** p for p in range(4)], 1)
    sinp = lawson(A, np.sin(2 * np.pi * t) / t, np.ones_like(t), 100).astype(np.float32)
    A = np.stack([t2 ** p for p in range(1, 5)], 1)
    cosp = lawson(A, np.cos(2 * np.pi * t) - 1, np.ones_like(t), 100).astype(np.float32)
    print("LOG_P", [hexf(v) for v in logp])
    print("SIN_S", [hexf(v) for v in sinp])
    print("COS_C", ["0x1p+0"] + [hexf(v) for v in cosp])


if __name__ == "__main__":
    main()
