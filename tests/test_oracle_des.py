"""The oracle's DES richness transfer INT_LC_LT_DES_t2
(int_lc_lt_des_t2.cuh:48-63) against the reference's own test
(test/int_lc_lt_des_t_test.test.cu: 720 expected values at the table knots,
compared there with epsilon 1e-3 and margin 1e-5) and against what the
mathematics fixes: exactness on bilinear tables between knots, the lambda_c
bin edges, and clamping outside the knots."""
import json
import os

import numpy as np

import oracle as O
from conftest import REF_DATA


def lclt_par(tables=None):
    """The packed parameter vector of the functor from the reference's data
    files (or from `tables`, four arrays [n_zt, n_lt], on the same knots)."""
    lt = np.loadtxt(os.path.join(REF_DATA, "int_lc_lt_des_t_lt_bins.txt"))
    zt = np.loadtxt(os.path.join(REF_DATA, "int_lc_lt_des_t_zt_bins.txt"))
    if tables is None:
        tables = [np.loadtxt(os.path.join(REF_DATA, f"int_lc_lt_des_t_lbd_bin{b}.txt"))
                  .reshape(-1) for b in range(4)]
    body = np.concatenate([np.asarray(t, dtype=float).reshape(-1) for t in tables])
    return np.ascontiguousarray(np.concatenate([[len(lt), len(zt)], lt, zt, body])), lt, zt


def f(par, lc, lt, zt):
    return O.lib.y3o_int_lc_lt_des(O.ptr(par), lc, lt, zt)


def test_int_lc_lt_des_reference_answers():
    with open(os.path.join(REF_DATA, "int_lc_lt_des_answers.json")) as fh:
        a = json.load(fh)
    par, _, _ = lclt_par()
    ans = np.array(a["answers"]).reshape(len(a["lc"]), len(a["zt"]), len(a["lt"]))
    n = 0
    for i, lc in enumerate(a["lc"]):
        for j, zt in enumerate(a["zt"]):
            for k, lt in enumerate(a["lt"]):
                got, exp = f(par, lc, lt, zt), ans[i, j, k]
                assert abs(got - exp) <= 1e-5 + 1e-3 * abs(exp), (lc, lt, zt, got, exp)
                n += 1
    assert n == 720


def test_int_lc_lt_des_bilinear_bins_and_clamp():
    """Tables c_b + a_b lt + d_b zt + e_b lt zt: reproduced exactly between
    knots; each lambda_c bin picks its own table (edges inclusive below);
    outside the knots the value is the edge's."""
    _, lt, zt = lclt_par()
    L, Z = np.meshgrid(lt, zt)  # [n_zt, n_lt], lt fastest
    coef = [(0.1, 0.01, 0.5, 0.002), (0.2, -0.003, 0.1, 0.01), (0.3, 0.004, -0.2, 0.003),
            (0.4, 0.002, 0.3, -0.001)]
    par, _, _ = lclt_par([c + a * L + d * Z + e * L * Z for c, a, d, e in coef])
    rng = np.random.default_rng(4)
    for lc, b in ((20.0, 0), (29.999, 0), (30.0, 1), (44.9, 1), (45.0, 2), (59.99, 2),
                  (60.0, 3), (150.0, 3), (19.99, 3), (3.0, 3)):
        c, a, d, e = coef[b]
        for _ in range(50):
            x, y = rng.uniform(lt[0], lt[-1]), rng.uniform(zt[0], zt[-1])
            assert abs(f(par, lc, x, y) - (c + a * x + d * y + e * x * y)) < 1e-12
        # clamped outside the knots
        x, y = lt[-1], zt[0]
        assert f(par, lc, 1e4, -1.0) == f(par, lc, x, y)
        assert abs(f(par, lc, 1e4, -1.0) - (c + a * x + d * y + e * x * y)) < 1e-12
