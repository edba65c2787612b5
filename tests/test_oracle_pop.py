"""Pins of the oracle's Costanzi-2026 P[X] operator integrands
(src/models/p_operator_gpu_t.cuh) and their pieces: the shifted-Poisson MOR
against the Poisson pmf, the photo-z table at its knots, the two-disk overlap
against a quadrature of chord lengths, the exclusion angle against the
triangle it solves, and identities between the four weights."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate, stats

import oracle as O

L = O.lib
MOR = [11.5, 12.8, 0.8, 0.25, 0.3, 0.45]   # log10 Mmin, log10 M1, alpha, sigma, eps, z_p


def pop_data(xi_const=None, bias_const=None, h0=0.7):
    """{h0, 0 x 4, MOR, n_chi, n_bm, n_bz, n_xr, n_xz, chi_z, chi_dc (Mpc),
    bias lnM, bias z, bias, xi r, xi z, xi} (y3g_des_upload's P[X] layout)."""
    zc = np.linspace(0.0, 1.5, 151)
    dc = np.array([L.y3o_comoving_distance(0.3, h0, z) / h0 for z in zc])
    bm = np.linspace(29.0, 37.0, 17)
    bz = np.linspace(0.0, 1.5, 7)
    B, Z = np.meshgrid(bm, bz)                      # [n_bz, n_bm], lnM fastest
    bias = (1 + 0.5 * (np.exp(B) / 1e14) ** 0.4) * (1 + Z) if bias_const is None \
        else np.full(B.shape, bias_const)
    xr = np.logspace(-1, 3.5, 60)
    xz = np.linspace(0.1, 1.0, 5)
    R, ZZ = np.meshgrid(xr, xz)                     # [n_xz, n_xr], r fastest
    xi = 5.0 * R ** -1.8 / (1 + ZZ) if xi_const is None else np.full(R.shape, xi_const)
    head = [h0, 0, 0, 0, 0] + MOR + [len(zc), len(bm), len(bz), len(xr), len(xz)]
    return np.ascontiguousarray(np.concatenate(
        [head, zc, dc, bm, bz, bias.reshape(-1), xr, xz, xi.reshape(-1)]))


def pop_integral(kind, zo=(0.3, 0.45), lam_bin=0):
    ig = O.Integral()
    ig.kind, ig.ndim = kind, 3
    ig.zo_lo, ig.zo_hi = zo
    ig.lo_lo = lam_bin
    ig.lo_hi = [25.0, 37.5, 52.5, 130.0][lam_bin]
    for a, (l, h) in enumerate([(0.0, 130.0), (0.1, 0.9), (29.9, 35.7)]):
        ig.lows[a], ig.highs[a] = l, h
    return ig


def test_shifted_poisson_is_poisson_pmf():
    q = np.ascontiguousarray(MOR)
    # at M = M1 and z = z_pivot, l_sat = 1: lam = 1 + sigma^2, mi = sigma^2
    lnM1 = MOR[1] * math.log(10.0)
    mi = MOR[3] ** 2
    lam = 1.0 + mi
    for k in range(0, 12):
        ltr = k + 1 - mi
        got = L.y3o_mor_shifted_poisson(O.ptr(q), ltr, lnM1, MOR[5])
        assert got == pytest.approx(stats.poisson.pmf(k, lam), rel=1e-12)
    # l_sat scales as ((M - Mmin) / (M1 - Mmin))^alpha ((1 + z) / (1 + z_p))^eps
    M = 10 ** MOR[0] + 7.0 * (10 ** MOR[1] - 10 ** MOR[0])
    ls = 7.0 ** MOR[2] * (1.6 / 1.45) ** MOR[4]
    mi = (ls * MOR[3]) ** 2
    got = L.y3o_mor_shifted_poisson(O.ptr(q), 4.0 + 1 - mi, math.log(M), 0.6)
    assert got == pytest.approx(stats.poisson.pmf(4, ls + mi), rel=1e-10)
    # zero at and below Mmin, and for ltr < 0
    assert L.y3o_mor_shifted_poisson(O.ptr(q), 3.0, MOR[0] * math.log(10.0) - 0.1, 0.3) == 0.0
    assert L.y3o_mor_shifted_poisson(O.ptr(q), -0.5, lnM1, 0.3) == 0.0


def test_photoz_table_knots():
    f = L.y3o_sigma_photoz_table
    assert f(0.1) == 0.0399897614683744 and f(0.0) == 0.0399897614683744
    assert f(0.9) == 0.11308048661154518 and f(2.0) == 0.11308048661154518
    z60 = 0.1 + 60 * (0.8 / 119)
    assert f(z60) == pytest.approx(0.08820842470461246, rel=1e-12)
    z61 = 0.1 + 61 * (0.8 / 119)
    assert f(0.5 * (z60 + z61)) == pytest.approx(0.5 * (0.08820842470461246 +
                                                        0.08798057371694759), rel=1e-12)


def _lens_fraction(th, lob, lt):
    """Area of the disk (radius lt, centre at distance th) inside the disk
    (radius lob, centre 0), by quadrature of vertical chord lengths, over
    pi lt^2."""
    def chord(x):
        ha = math.sqrt(max(lt * lt - (x - th) ** 2, 0.0))
        hb = math.sqrt(max(lob * lob - x * x, 0.0)) if abs(x) <= lob else 0.0
        return 2 * min(ha, hb)
    pts = sorted(set([th - lt, th + lt, -lob, lob]))
    a, b = th - lt, th + lt
    v = integrate.quad(chord, a, b, points=[p for p in pts if a < p < b], limit=400,
                       epsabs=0, epsrel=1e-11)[0]
    return v / (math.pi * lt * lt)


def test_area_overlap_is_the_lens_area():
    for lob, lt in [(1.0, 0.3), (0.4, 1.1), (1.0, 1.0), (2.0, 0.05)]:
        # (24 points: none at the exact tangencies th = lob + lt, |lob - lt|,
        # where the closed form's cancellation leaves ~1e-4 of the disk)
        for th in np.linspace(0.0, 1.1 * (lob + lt), 24):
            got = L.y3o_pop_area_overlap(th, lob, lt)
            assert got == pytest.approx(_lens_fraction(th, lob, lt), abs=1e-9), (lob, lt, th)


def test_theta_excl_solves_the_triangle():
    for cz, co, R in [(1000.0, 1200.0, 300.0), (800.0, 790.0, 12.0), (500.0, 520.0, 21.5)]:
        th = L.y3o_pop_theta_excl(cz, co, R)
        d = math.hypot(cz - co * math.cos(th), co * math.sin(th))
        assert d == pytest.approx(R, rel=1e-9)
    # no triangle (R < |chi_z - chi_o|) or a degenerate one: the floor 1e-6
    assert L.y3o_pop_theta_excl(500.0, 520.0, 1.5) == 1e-6
    assert L.y3o_pop_theta_excl(1000.0, 1000.0, 0.0) == 1e-6


def _eval(kind, des, ig, x):
    p = O.default_params()
    t = O.OracleTables(p, O.spec_wide())
    return L.y3o_integrand_des(C.byref(p), C.byref(t.t), O.ptr(des), C.byref(ig),
                               O.ptr(np.ascontiguousarray(x)))


def test_weight_identities():
    des = pop_data()
    rng = np.random.default_rng(2)
    n_nonzero = 0
    for _ in range(60):
        x = [rng.uniform(1.0, 25.0), rng.uniform(0.3, 0.45), rng.uniform(32.0, 34.5)]
        v = {k: _eval(k, des, pop_integral(k), x) for k in (8, 9, 10, 11)}
        assert v[9] + v[11] == pytest.approx(v[10], rel=1e-12, abs=1e-300)
        n_nonzero += v[10] != 0.0
    assert n_nonzero > 20
    # constant xi_NL (c) and bias (b): I2 = b c P1
    dc = pop_data(xi_const=0.37, bias_const=1.9)
    for _ in range(20):
        x = [rng.uniform(1.0, 25.0), rng.uniform(0.3, 0.45), rng.uniform(32.0, 34.5)]
        p1 = _eval(8, dc, pop_integral(8), x)
        assert _eval(10, dc, pop_integral(10), x) == pytest.approx(1.9 * 0.37 * p1, rel=1e-12,
                                                                  abs=1e-300)


def test_integrand_support():
    des = pop_data()
    ig = pop_integral(8)
    # lt beyond the bin centre, lt <= 0, zt beyond sigma_z of the bin centre
    assert _eval(8, des, ig, [26.0, 0.375, 33.0]) == 0.0
    assert _eval(8, des, ig, [0.0, 0.375, 33.0]) == 0.0
    s = L.y3o_sigma_photoz_table(0.6)
    assert _eval(8, des, ig, [10.0, 0.375 + 1.01 * 0.2 + s, 33.0]) == 0.0
    assert _eval(8, des, ig, [10.0, 0.375, 33.0]) > 0.0


def test_angle_sum_against_adaptive_quadrature():
    """The fixed 2 x 5-point Gauss-Legendre angle sum of the integrand,
    recovered as integrand / (w_z lt hmf dV/dz MOR b), against adaptive
    quadrature of the same angle integrand (the overlap's kinks bound the
    agreement: a few per cent)."""
    p = O.default_params()
    t = O.OracleTables(p, O.spec_wide())
    des = pop_data(bias_const=1.0)
    q = des[11:]
    nc = int(q[0])
    zc, dc = q[5:5 + nc], q[5 + nc:5 + 2 * nc]
    nbm, nbz, nxr, nxz = (int(v) for v in q[1:5])
    o = 5 + 2 * nc + nbm + nbz + nbm * nbz
    xr = np.ascontiguousarray(q[o:o + nxr])
    xz = np.ascontiguousarray(q[o + nxr:o + nxr + nxz])
    xt = np.ascontiguousarray(q[o + nxr + nxz:])
    for kind in (8, 10):
        ig = pop_integral(kind)
        for lt, zt, lnM in [(12.0, 0.37, 33.0), (20.0, 0.4, 33.5)]:
            zob = 0.375
            w = 1 - ((zt - zob) / L.y3o_sigma_photoz_table(zt)) ** 2
            common = (L.y3o_hmf(C.byref(p), C.byref(t.t), lnM, zt) *
                      L.y3o_dvdz(C.byref(p), C.byref(t.t), zt) *
                      L.y3o_mor_shifted_poisson(O.ptr(np.ascontiguousarray(MOR)), lt, lnM, zt))
            th_sum = _eval(kind, des, ig, [lt, zt, lnM]) / (w * lt * common)
            chi_o = np.interp(zob, zc, dc) * 0.7
            chi_z = np.interp(zt, zc, dc) * 0.7
            rl = (25.0 / 100) ** 0.2
            tlob = rl * (1 + zob) / chi_o
            tex = L.y3o_pop_theta_excl(chi_z, chi_o, rl * (1 + zob))
            tlt = (lt / 100) ** 0.2 * (1 + zt) / chi_z
            f = lambda th: L.y3o_pop_theta_term(kind, th, tlob, tlt, chi_z, chi_o, zob,
                                                O.ptr(xr), nxr, O.ptr(xz), nxz, O.ptr(xt))
            pts = [v for v in (abs(tlob - tlt), tlob + tlt, tlob) if tex < v < 2 * tlob]
            ref = integrate.quad(f, tex, 2 * tlob, points=pts, limit=400)[0]
            assert th_sum == pytest.approx(ref, rel=5e-2), (kind, lt, zt, th_sum, ref)
