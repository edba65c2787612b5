"""Table of the NFW Delta-Sigma shape g(x) for the FP32 device path.

g(x) = Delta-Sigma / (rho_s r_s) of an NFW halo at x = R / r_s (Wright &
Brainerd 2000; the FP64 device and CPU forms evaluate it in closed form).  The
FP32 path interpolates v(s) = log2(g(x) (1 + x^2)) in s = log2(x^2), a smooth
function of bounded derivatives (the factor 1 + x^2 removes the x^-2 decline,
so that v stays within +-6), by cubic Hermite interpolation on a uniform grid
of s: knots every 1/4 in s on [-40, 32] (x from 1e-6 to 6.5e4), each with its
value and its slope times the knot spacing, both from the closed form in FP64
(slope by central difference).  Largest relative error of the FP32
evaluation on x in [1e-5, 3e4]: 6.8e-7.

    python y3_cluster_cpp_amd/csrc/fit_nfw_table.py   # rewrites y3g_nfw_table.inc
"""
import os

import numpy as np

S0, S1, STEP = -40.0, 32.0, 0.25


def g64(x):
    """Closed form, FP64: 4 m(x) - 2 h(x) with F, h and the mass term m."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    out = np.empty_like(x)
    for i, xx in enumerate(x):
        t = xx * xx - 1.0
        if abs(t) < 0.1:
            h = sum((-t) ** (k - 1) / (2 * k + 1) for k in range(1, 25))
            F = 1.0 - t * h
        elif t < 0:
            s = np.sqrt(-t)
            F = np.log((1 + s) / xx) / s
            h = (1 - F) / t
        else:
            s = np.sqrt(t)
            F = np.arctan(s) / s
            h = (1 - F) / t
        if xx < 0.9:
            s = np.sqrt(1 - xx * xx)
            q = xx * xx / (2 * (1 + s))
            phi = -np.log1p(-q) / q
            m = (np.log(2 / xx) / (1 + s) - 0.5 * phi / (1 + s)) / s
        else:
            m = (np.log(0.5 * xx) + F) / (xx * xx)
        out[i] = 4 * m - 2 * h
    return out


def v_of(s):
    x = np.sqrt(2.0 ** np.asarray(s, dtype=np.float64))
    return np.log2(g64(x) * (1.0 + x * x))


def knots():
    s = np.arange(S0, S1 + STEP / 2, STEP)
    e = 1e-5
    return v_of(s), (v_of(s + e) - v_of(s - e)) / (2 * e) * STEP


def eval_f32(x2, v, d):
    f = np.float32
    x2 = f(x2)
    u = (np.log2(x2) - f(S0)) * f(1 / STEP)
    u = min(max(u, f(0)), f(len(v) - 1) - f(1e-3))
    i = int(u)
    t = f(u - i)
    t2 = t * t
    t3 = t2 * t
    lg = ((f(2) * t3 - f(3) * t2 + f(1)) * f(v[i]) + (t3 - f(2) * t2 + t) * f(d[i]) +
          (f(-2) * t3 + f(3) * t2) * f(v[i + 1]) + (t3 - t2) * f(d[i + 1]))
    return np.exp2(lg) * (f(1) / (f(1) + x2))


if __name__ == "__main__":
    v, d = knots()
    xs = np.concatenate([np.geomspace(1e-5, 3e4, 4000), np.linspace(0.9, 1.1, 401)])
    err = np.abs(np.array([eval_f32(x * x, v, d) for x in xs]) / g64(xs) - 1)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "y3g_nfw_table.inc")
    with open(path, "w") as fo:
        fo.write("// Generated by fit_nfw_table.py: cubic Hermite knots of\n"
                 "// v(s) = log2(g(x) (1 + x^2)), s = log2(x^2), for the FP32 NFW\n"
                 f"// Delta-Sigma shape; largest relative error {err.max():.2e}.\n")
        fo.write(f"constexpr float kNfwS0 = {S0:.1f}f;\nconstexpr float kNfwInvStep = {1 / STEP:.1f}f;\n"
                 f"constexpr int kNfwN = {len(v)};\n")
        fo.write("struct NfwKnot {\n  float v, d;  // value, slope x knot spacing\n};\n")
        fo.write("static __constant__ NfwKnot kNfwTab[kNfwN] = {\n")
        for a, b in zip(v, d):
            fo.write(f"  {{{np.float32(a)!r}f, {np.float32(b)!r}f}},\n".replace("np.float32(", "").replace(")f", "f"))
        fo.write("};\n")
    print("wrote", path, "entries", len(v), "max rel err (FP32)", err.max())
