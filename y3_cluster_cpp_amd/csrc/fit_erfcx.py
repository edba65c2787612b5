"""Coefficients of the FP32 erfcx used by the device EMG density.

Method: M. M. Shepherd and J. G. Laframboise, "Chebyshev approximation of
(1 + 2x) exp(x^2) erfc x in 0 <= x < infinity", Math. Comp. 36 (1981)
249-253: with t = (x - K) / (x + K), (1 + 2x) erfcx(x) is a smooth function
of t on [-1, 1).  Here K = 2 and a degree-10 polynomial in t, fitted by least
squares at 4000 Chebyshev points of t against scipy's erfcx and converted to
the power basis for Horner evaluation.  Prints the coefficients (constant
term first) and the largest relative error of the FP32 evaluation on
x in [0, 1e7].

    python y3_cluster_cpp_amd/csrc/fit_erfcx.py
"""
import numpy as np
import scipy.special as sp
from numpy.polynomial import chebyshev as Ch

K, DEG, N = 2.0, 10, 4000


def coefficients():
    t = np.cos(np.pi * (np.arange(N) + 0.5) / N)
    x = K * (1 + t) / (1 - t)
    c = Ch.chebfit(t, (1 + 2 * x) * sp.erfcx(x), DEG)
    return Ch.cheb2poly(c)


def erfcx_f32(x, pc):
    f = np.float32
    x = np.asarray(x, dtype=f)
    t = (x - f(K)) * (f(1) / (x + f(K)))
    acc = np.full_like(t, f(pc[-1]))
    for a in pc[-2::-1]:
        acc = acc * t + f(a)
    return acc * (f(1) / (f(1) + f(2) * x))


if __name__ == "__main__":
    pc = coefficients()
    for k, a in enumerate(pc):
        print(f"c{k} = {np.float32(a)!r:>16}  ({a:.17g})")
    xs = np.concatenate([np.linspace(0, 10, 200001), np.geomspace(10, 1e7, 20000)])
    print("max rel err (FP32):", np.max(np.abs(erfcx_f32(xs, pc) / sp.erfcx(xs) - 1)))
