"""The hat-smoothed HMF of the oracle (y3o_hmf_coefs, y3o_hmf_smooth) and the
smoothing identity behind it: the integral of the bilinear table against a
smooth factor equals that of the cubic spline convolved with the one-cell hat
kernel, a tensor quintic B-spline over the spline's coefficients.

Pinned independently of the oracle's formulas: the spline interpolates the
table (B3 values 1/6, 4/6, 1/6 at the knots) with not-a-knot ends and
reproduces cubics, whose hat convolution is p + h^2 p''/12; the quintic
weights against scipy's B-spline basis; exact reproduction of a bilinear
surface; the identity against quadrature of the bilinear table itself,
interior and at the table's edges; and the cubature of the smoothed integrand
against the golden value of the bilinear one.
"""
import ctypes as C

import numpy as np
import pytest
from scipy.interpolate import BSpline

import oracle as O

L = O.lib


def coefs_1d(y):
    y = np.ascontiguousarray(y, dtype=np.float64)
    c = np.zeros(len(y) + 4)
    L.y3o_spline_coefs(O.ptr(y), len(y), 1, O.ptr(c), 1)
    return c  # index k + 2, k = -2 .. n + 1


def smooth_1d(c, x0, h, x):
    """sum_k c_k B5((x - x_k) / h) through the oracle's quintic weights."""
    n = len(c) - 4
    u = min(max((x - x0) / h, 0.0), n - 1.0)
    i = min(int(u), n - 2)
    w = np.zeros(6)
    L.y3o_b5_weights(u - i, O.ptr(w))
    return float(np.dot(w, c[i:i + 6]))


def test_spline_interpolates_with_not_a_knot_ends():
    rng = np.random.default_rng(3)
    y = np.cumsum(rng.normal(size=40)) + 10.0
    c = coefs_1d(y)
    k = np.arange(40) + 2
    assert np.allclose((c[k - 1] + 4 * c[k] + c[k + 1]) / 6.0, y, rtol=0, atol=1e-12)
    # not-a-knot: the third derivative is continuous at x_1 and x_{n-2}
    # (fourth differences of the coefficients vanish there)
    d4 = c[:-4] - 4 * c[1:-3] + 6 * c[2:-2] - 4 * c[3:-1] + c[4:]
    assert abs(d4[1]) < 1e-11 * abs(y).max() and abs(d4[-2]) < 1e-11 * abs(y).max()


def test_cubic_data_and_its_hat_convolution():
    """A cubic's not-a-knot spline is the cubic; its hat convolution is
    p + h^2 p'' / 12 (the one-cell hat has second moment h^2 / 6)."""
    n, x0, h = 30, 1.0, 0.1
    xs = x0 + h * np.arange(n)
    p = lambda x: 0.3 - 1.2 * x + 0.7 * x ** 2 - 0.25 * x ** 3  # noqa: E731
    p2 = lambda x: 1.4 - 1.5 * x  # noqa: E731
    c = coefs_1d(p(xs))
    for x in np.linspace(xs[0], xs[-1], 37):
        assert abs(smooth_1d(c, x0, h, x) - (p(x) + h * h * p2(x) / 12.0)) < 1e-12


def test_b5_weights_against_scipy_basis():
    B5 = BSpline.basis_element(np.arange(-3.0, 4.0), extrapolate=False)
    for f in [0.0, 0.17, 0.5, 0.83, 0.999]:
        w = np.zeros(6)
        L.y3o_b5_weights(f, O.ptr(w))
        ref = [float(B5(f + o)) if -3 <= f + o <= 3 else 0.0 for o in (2, 1, 0, -1, -2, -3)]
        assert np.allclose(w, ref, rtol=0, atol=1e-14)
        assert abs(w.sum() - 1.0) < 1e-14


def test_smoothed_table_reproduces_a_bilinear_surface():
    # linear data: the spline is the line itself and the hat
    # convolution of a bilinear surface is the surface
    p = O.default_params()
    s = O.spec_wide()
    t = O.OracleTables(p, s)
    a, b, cc, d = 2.0, 0.3, -1.1, 0.05
    lnm = t.lnm
    z = t.z
    tab = (a + b * lnm[None, :] + cc * z[:, None] + d * lnm[None, :] * z[:, None]).ravel()
    hm = np.ctypeslib.as_array(t.t.hmf, (s.n_z * s.n_lnm,))
    hm[:] = tab
    L.y3o_hmf_coefs(t.t.hmf, s.n_lnm, s.n_z, t.t.hmfc)
    rng = np.random.default_rng(5)
    for _ in range(50):
        x = rng.uniform(s.lnm_lo, s.lnm_hi)
        zz = rng.uniform(s.z_lo, s.z_hi)
        exact = a + b * x + cc * zz + d * x * zz
        assert abs(L.y3o_hmf_smooth(C.byref(p), C.byref(t.t), x, zz) - exact) < 1e-12 * abs(exact) + 1e-13
        assert abs(L.y3o_hmf(C.byref(p), C.byref(t.t), x, zz) - exact) < 1e-12 * abs(exact) + 1e-13


def _hmf_like(x):
    sig = 0.8 * np.exp(-0.25 * (x - 32.236))
    return np.exp(-1.2 / sig ** 2) * sig ** -1.5


@pytest.mark.parametrize("mu,sg", [(33.3, 0.15), (33.3, 0.05), (35.5, 0.3), (35.69, 0.5)])
def test_identity_one_axis(mu, sg):
    """Int lerp(y) g  vs  Int (spline * hat) g  on the HMF's lnM grid, with g
    narrow (1.7 cells) or weighted at the table's upper edge."""
    n = 200
    x0, x1 = 29.933606208922594, 35.690068941407709
    h = (x1 - x0) / (n - 1)
    xs = x0 + h * np.arange(n)
    y = _hmf_like(xs)
    c = coefs_1d(y)
    gx, gw = np.polynomial.legendre.leggauss(10)
    g = lambda x: np.exp(-0.5 * ((x - mu) / sg) ** 2)  # noqa: E731
    lin = smo = 0.0
    for i in range(n - 1):
        for sub in range(4):  # both sides smooth on each quarter cell
            a = xs[i] + sub * h / 4
            b = a + h / 4
            xx = 0.5 * (a + b) + 0.5 * (b - a) * gx
            ww = 0.5 * (b - a) * gw
            lin += np.sum(ww * np.interp(xx, xs, y) * g(xx))
            smo += np.sum(ww * np.array([smooth_1d(c, x0, h, v) for v in xx]) * g(xx))
    assert abs(smo / lin - 1) < 2e-7
    # and the bilinear form differs from the smooth function it samples by
    # far more: the identity is not the statement that both approximate it
    true = sum(np.sum(0.5 * (b - a) * gw * _hmf_like(0.5 * (a + b) + 0.5 * (b - a) * gx) *
                      g(0.5 * (a + b) + 0.5 * (b - a) * gx)) for a, b in zip(xs[:-1], xs[1:]))
    assert abs(lin / true - 1) > 20 * abs(smo / lin - 1)


def test_identity_two_axes_on_the_table():
    """The 2-D identity on a config-3 HMF table against a ridge in (lnM, z)
    like the MOR's: per-cell Gauss-Legendre of the bilinear table vs of the
    smoothed one."""
    p = O.default_params(omega_m=0.35, hmf_amp=0.75)
    s = O.spec_wide()
    t = O.OracleTables(p, s)
    gx, gw = np.polynomial.legendre.leggauss(6)
    g = lambda m, z: np.exp(-0.5 * ((m - 33.0 - 0.8 * (z - 0.4)) / 0.2) ** 2 -  # noqa: E731
                            0.5 * ((z - 0.4) / 0.06) ** 2)
    lin = smo = 0.0
    mk, zk = t.lnm, t.z
    for j in range(s.n_z - 1):
        z0, z1 = zk[j], zk[j + 1]
        if not (0.2 < 0.5 * (z0 + z1) < 0.6):
            continue
        for i in range(s.n_lnm - 1):
            m0, m1 = mk[i], mk[i + 1]
            if not (31.8 < 0.5 * (m0 + m1) < 34.2):
                continue
            for a in range(6):
                for b in range(6):
                    z = 0.5 * (z0 + z1) + 0.5 * (z1 - z0) * gx[a]
                    m = 0.5 * (m0 + m1) + 0.5 * (m1 - m0) * gx[b]
                    w = 0.25 * (z1 - z0) * (m1 - m0) * gw[a] * gw[b] * g(m, z)
                    lin += w * L.y3o_hmf(C.byref(p), C.byref(t.t), m, z)
                    smo += w * L.y3o_hmf_smooth(C.byref(p), C.byref(t.t), m, z)
    assert abs(smo / lin - 1) < 2e-7


def test_smoothed_cubature_meets_the_bilinear_golden():
    """Config 0 (3-D count over exactly the table's z range): the cubature of
    the smoothed integrand against the golden value of the bilinear one, with
    the reported error (which carries the 1e-6 smoothing bound) covering it."""
    p = O.default_params(f_prj=0.0)
    s = O.spec_config0()
    t = O.OracleTables(p, s)
    gold = O.golden(p, t, O.COUNT3D, [20., 30.])[0][0, 0]
    ig = O.make_integral(p, s, O.COUNT3D)
    r = O.cubature(p, t, ig, O.cub_opts(eps_rel=1e-5, max_regions=100000, max_eval=10**9))
    assert r.status == 0
    assert abs(r.value - gold) <= r.error
    assert abs(r.value / gold - 1) < 1e-5
    assert L.y3o_smooth_bound(O.COUNT3D, 1) == 1e-6 and L.y3o_smooth_bound(O.COUNT3D, 0) == 0.0
    assert L.y3o_smooth_bound(3, 1) == 0.0  # real-data kinds read the bilinear table
