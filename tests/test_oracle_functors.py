"""Pins of the oracle's model functors: the reference's own known-answer
fixtures (same tolerances as its Catch2 tests) and closed forms."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate

import oracle as O
from conftest import load_columns

L = O.lib


def approx(actual, expected, epsrel, epsabs=1e-12):
    # Catch2 v2 Approx(expected).epsilon(epsrel).margin(epsabs)
    return abs(actual - expected) <= max(epsabs, epsrel * abs(expected))


def test_roffset_gamma_fixture():
    # test/roffset_t.test.cc:40-46, tau = 0.15, epsilon 1e-6
    d = load_columns("test_Rmis.txt", 2)
    for r, p in d:
        assert approx(L.y3o_roffset_gamma(r, 0.15), p, 1e-6)


def test_lo_lc_fixture():
    # test/lo_lc_t.test.cc:47-58, LO_LC_t(1.66, 0.26, 1.43, 1.0), checks f*lc
    d = load_columns("test_lolc_t.txt", 2)
    assert len(d) > 700
    for lo, lc, R, y in d:
        f = L.y3o_lo_lc(lo, lc, R, 1.66, 0.26, 1.43, 1.0)
        assert approx(f * lc, y, 1e-6)


def test_int_zo_zt_des_fixture():
    # test/int_zo_zt_des_t.test.cc:49-58
    d = load_columns("test_int_zo_zt_des.txt", 2)
    for zmin, zmax, zt, sig, val in d:
        assert approx(L.y3o_sigma_photoz_des(zt), sig, 1e-6)
        assert approx(L.y3o_int_zo_zt_des(zmin, zmax, zt), val, 1e-6)


def test_zo_zt_des_fixture():
    # test/zo_zt_des_t.test.cc:45-53
    d = load_columns("test_zo_zt_des.txt", 2)
    for zo, zt, sig, prob in d:
        assert approx(L.y3o_sigma_photoz_des(zt), sig, 1e-6)
        assert approx(L.y3o_zo_zt_des(zo, zt), prob, 1e-6)


def test_dv_do_dz_fixture():
    # test/dv_do_dz_t.test.cc:34-52: D_A table, EZ(0.3, 0.7, 0), h = 0.7, 1e-3
    z = np.ascontiguousarray(load_columns("z_da_test.txt")[:, 0])
    da = np.ascontiguousarray(load_columns("d_a_test.txt")[:, 0])
    d = load_columns("test_dv_do_dz.txt")
    for zi, yi in d[1:]:
        v = L.y3o_dv_do_dz(zi, O.ptr(z), O.ptr(da), len(z), 0.3, 0.7, 0.0, 0.7)
        assert approx(v, yi, 1e-3)


def test_comoving_distance_matches_da_fixture():
    # the reference's D_A table (Mpc, Omega_m = 0.3, h = 0.7) against the
    # oracle's flat-LCDM quadrature: D_A = chi / h / (1+z)
    z = load_columns("z_da_test.txt")[:, 0]
    da = load_columns("d_a_test.txt")[:, 0]
    for zi, di in list(zip(z, da))[1::25]:
        mine = L.y3o_comoving_distance(0.3, 0.7, zi) / 0.7 / (1 + zi)
        assert abs(mine / di - 1) < 2e-4


def test_comoving_distance_closed_form():
    # Omega_m = 1: chi = 2 DH (1 - 1/sqrt(1+z))
    for z in (0.1, 0.5, 1.0, 2.0):
        exact = 2 * 2997.92458 * (1 - 1 / math.sqrt(1 + z))
        assert abs(L.y3o_comoving_distance(1.0, 0.7, z) / exact - 1) < 1e-13


def _grid_fxy(xs, ys, f):
    zs = np.array([[f(x, y) for x in xs] for y in ys])  # x-fastest rows
    return np.ascontiguousarray(xs, dtype=float), np.ascontiguousarray(ys, dtype=float), \
        np.ascontiguousarray(zs.reshape(-1))


def test_interp2d_cases():
    # test/interp_2d.test.cc: f = 3xy + 2x + 4y on x {1,2,3}, y {4,5}
    f = lambda x, y: 3 * x * y + 2 * x + 4 * y
    xs, ys, zs = _grid_fxy([1., 2., 3.], [4., 5.], f)
    ev = lambda x, y: L.y3o_interp2d_clamp(O.ptr(xs), 3, O.ptr(ys), 2, O.ptr(zs), x, y)
    out = C.c_double()
    assert L.y3o_interp2d(O.ptr(xs), 3, O.ptr(ys), 2, O.ptr(zs), 2.5, 4.5, C.byref(out)) == 0
    assert out.value == 56.75
    assert ev(0, 4.5) == ev(1, 4.5) and ev(4, 4.5) == ev(3, 4.5)
    assert ev(2, 3) == ev(2, 4) and ev(2, 5.5) == ev(2, 5)
    assert ev(0, 0) == ev(1, 4) and ev(4, 3) == ev(3, 4)
    assert ev(0, 6) == ev(1, 5) and ev(4, 6) == ev(3, 5)
    for x in (1., 2., 3.):
        for y in (4., 5.):
            assert ev(x, y) == f(x, y)
    # out of range is an error for the unclamped call
    assert L.y3o_interp2d(O.ptr(xs), 3, O.ptr(ys), 2, O.ptr(zs), 0.5, 4.5, C.byref(out)) != 0
    # exact on bilinear functions
    g = lambda x, y: 2 * x + 3 * y - 5
    xs, ys, zs = _grid_fxy([1., 2., 3.], [1., 2., 3., 4.], g)
    rng = np.random.default_rng(1)
    for x, y in rng.uniform([1, 1], [3, 4], size=(50, 2)):
        v = L.y3o_interp2d_clamp(O.ptr(xs), 3, O.ptr(ys), 4, O.ptr(zs), x, y)
        assert abs(v - g(x, y)) < 1e-13


def test_interp1d_cases():
    xs = np.ascontiguousarray([0., 1., 3.])
    ys = np.ascontiguousarray([1., 3., 7.])
    out = C.c_double()
    assert L.y3o_interp1d(O.ptr(xs), O.ptr(ys), 3, 2.0, C.byref(out)) == 0 and out.value == 5.0
    assert L.y3o_interp1d(O.ptr(xs), O.ptr(ys), 3, 3.5, C.byref(out)) != 0
    assert L.y3o_interp1d_clamp(O.ptr(xs), O.ptr(ys), 3, 9.0) == 7.0


def test_emg_cdf_derivative_and_normalisation():
    for (mu, sig, tau) in [(25., 5., 0.4), (100., 10., 0.1), (5., 2.2, 2.0)]:
        tot, _ = integrate.quad(lambda x: L.y3o_emg_pdf(x, mu, sig, tau), -np.inf, np.inf,
                                epsabs=0, epsrel=1e-12, limit=400)
        assert abs(tot - 1) < 1e-9
        for x in np.linspace(mu - 3 * sig, mu + 5 * sig, 9):
            h = 1e-4 * sig
            dnum = (L.y3o_emg_cdf(x + h, mu, sig, tau) - L.y3o_emg_cdf(x - h, mu, sig, tau)) / (2 * h)
            assert abs(dnum - L.y3o_emg_pdf(x, mu, sig, tau)) < 1e-7
        assert abs(L.y3o_emg_cdf(mu + 60 * sig, mu, sig, tau) - 1) < 1e-12


def test_lo_lt_bin_is_integral_of_pdf():
    p = O.default_params()
    for lt in (8., 15., 25., 40., 90.):
        v, _ = integrate.quad(lambda lo: L.y3o_lo_lt_pdf(C.byref(p), lo, lt), 20, 30,
                              epsabs=0, epsrel=1e-13, limit=200)
        b = L.y3o_lo_lt_bin(C.byref(p), 20., 30., lt)
        assert abs(b - v) <= 1e-12 * max(v, 1e-300) + 1e-16


def test_mor_normalised():
    p = O.default_params()
    for lnM, zt in [(32.0, 0.3), (33.5, 0.5), (34.5, 0.2)]:
        v, _ = integrate.quad(lambda lt: L.y3o_mor_lognormal(C.byref(p), lt, lnM, zt), 0, np.inf,
                              epsabs=0, epsrel=1e-12, limit=400)
        assert abs(v - 1) < 1e-9


def test_lc_lt_eq15_normalisation():
    # test/lc_lt_probability_unity.test.cc pattern: int dlc = 1 (+-1e-2)
    for (lt, fmsk, fprj) in [(20., 0.0, 0.2), (50., 0.1, 0.2), (80., 0.3, 0.4)]:
        mu, sig, tau = lt, math.sqrt(lt), 1.0 / (0.1 * lt)
        v, _ = integrate.quad(lambda lc: L.y3o_lc_lt_eq15(lc, lt, mu, sig, tau, fmsk, fprj),
                              -200, 1000, epsabs=0, epsrel=1e-10, limit=800, points=[lt])
        assert abs(v - 1) < 1e-2
    # fmsk = 0 is (1 - fprj) Gauss + fprj EMG
    lt = 30.
    for lc in (20., 30., 45.):
        a = L.y3o_lc_lt_eq15(lc, lt, lt, math.sqrt(lt), 1 / (3.0), 0.0, 0.2)
        b = 0.8 * L.y3o_gaussian(lc, lt, math.sqrt(lt)) + 0.2 * L.y3o_emg_pdf(lc, lt, math.sqrt(lt), 1 / 3.0)
        assert abs(a - b) < 1e-12 * abs(b)


def test_nfw_sigma_line_of_sight():
    # Sigma(x) / (rho_s r_s) = 2 h(x) against 2 int_0^inf rho(sqrt(x^2+s^2)) ds
    rho = lambda r: 1.0 / (r * (1 + r) ** 2)
    for x in (0.05, 0.5, 0.99, 1.0, 1.01, 2.0, 10.0):
        v, _ = integrate.quad(lambda s: 2 * rho(math.sqrt(x * x + s * s)), 0, np.inf,
                              epsabs=0, epsrel=1e-12, limit=400)
        assert abs(L.y3o_nfw_sigma_unit(x) / v - 1) < 1e-9


def test_nfw_delta_sigma_mean_minus_local():
    for x in (0.05, 0.3, 0.95, 1.0, 1.05, 3.0, 20.0):
        inner, _ = integrate.quad(lambda r: L.y3o_nfw_sigma_unit(r) * r, 0, x,
                                  epsabs=0, epsrel=1e-12, limit=400)
        mean = 2 * inner / (x * x)
        assert abs(L.y3o_nfw_g(x) - (mean - L.y3o_nfw_sigma_unit(x))) < 1e-9 * mean
    # Wright & Brainerd value at x = 1
    assert abs(L.y3o_nfw_g(1.0) - (10 / 3 + 4 * math.log(0.5))) < 1e-14


# cluster_toolkit values of test/nfw_dsigma_mis.test.cu:32-49 (c = 4,
# rho = 0.3 rho_crit, lnM = 32.2361913, gamma kernel), tolerance 5e-3
CT_R = [0.1, 0.16681005, 0.27825594, 0.46415888, 0.77426368, 1.29154967,
        2.15443469, 3.59381366, 5.9948425, 10.]
CT_RMIS = [0.1, 0.5, 2.0]
CT_VALUES = [12.3685326, 0.3899747, 0.0104446, 22.7491859, 0.9300333, 0.0264995,
             33.3082693, 2.0334094, 0.0641354, 36.0917619, 3.9092169, 0.1455426,
             28.3882168, 6.2393766, 0.3031914, 17.5178459, 7.6804544, 0.5638920,
             9.4010319, 6.7788772, 0.9002713, 4.6358405, 4.2448975, 1.1650680,
             2.1556775, 2.1118245, 1.1349319, 0.9599990, 0.9545601, 0.7860161]


@pytest.mark.parametrize("j", range(3))
def test_dsigma_mis_cluster_toolkit(j):
    rho = 0.3 * 2.77533742639e11
    for i, R in enumerate(CT_R):
        exp = CT_VALUES[j + 3 * i]
        got = L.y3o_dsigma_mis_gamma(R, CT_RMIS[j], 32.2361913, rho, 4.0, 1e-8)
        assert abs(got / exp - 1) < 5e-3, (R, CT_RMIS[j], got, exp)


def test_dsigma_offset_limits():
    rho = 0.3 * 2.77533742639e11
    # no offset: tangential Delta-Sigma is the centred one
    for R in (0.2, 1.0, 5.0):
        a = L.y3o_dsigma_offset(R, 0.0, 1.3, 32.5, rho, 4.0)
        b = L.y3o_nfw_dsigma(R, 32.5, rho, 4.0)
        assert abs(a / b - 1) < 1e-14
    # ring average with a small offset tends to the centred value
    R = 1.0
    ring = L.y3o_dsigma_mis_ring(R, 1e-4, 32.5, rho, 4.0, 1e-12)
    assert abs(ring / L.y3o_nfw_dsigma(R, 32.5, rho, 4.0) - 1) < 1e-6


def test_hmf_amplitude_factor():
    """HMF_t's s (lnM log10 e - 13.8124426028) + q (hmf_t.cuh:54-60): at
    log10 M = 13.8124426028 the factor is q whatever s, and it is linear in
    lnM with slope s log10 e."""
    p0 = O.default_params()
    t = O.OracleTables(p0, O.spec_wide())
    lnm_piv = 13.8124426028 / 0.4342944819
    z = 0.4
    base = O.lib.y3o_hmf(C.byref(p0), C.byref(t.t), lnm_piv, z)
    p1 = O.default_params(hmf_s=0.3, hmf_q=0.8)
    assert O.lib.y3o_hmf(C.byref(p1), C.byref(t.t), lnm_piv, z) == pytest.approx(0.8 * base, rel=1e-12)
    lnm = lnm_piv + 1.0
    b1 = O.lib.y3o_hmf(C.byref(p0), C.byref(t.t), lnm, z)
    assert O.lib.y3o_hmf(C.byref(p1), C.byref(t.t), lnm, z) == pytest.approx(
        (0.8 + 0.3 * 0.4342944819) * b1, rel=1e-12)


# ---- pins of the synthetic-config functors: each value is rebuilt here from
# ---- the configuration's own definition with independent arithmetic

def test_hmf_formula_hand_values():
    """dn/dlnM = A (sigma^-a + 1) exp(-c / sigma^2) with
    sigma = s_8 (M / 1e14)^-0.25 / (1 + z)^0.8, A = 0.2, a = 1.5, c = 1.2,
    evaluated from M and z directly (powers of ten, no logarithms)."""
    def hand(M, z, s8):
        sig = s8 * (M / 1e14) ** -0.25 / (1 + z) ** 0.8
        return 0.2 * (sig ** -1.5 + 1) * math.exp(-1.2 / sig ** 2)
    # at M = 1e14, z = 0 sigma is the amplitude itself: 0.2 (0.8^-1.5 + 1) e^{-1.875}
    p = O.default_params()
    assert L.y3o_hmf_formula(C.byref(p), math.log(1e14), 0.0) == pytest.approx(
        0.2 * (0.8 ** -1.5 + 1) * math.exp(-1.2 / 0.64), rel=1e-14)
    for M, z, s8 in [(1e13, 0.2, 0.8), (10 ** 15.5, 0.35, 0.8), (3e14, 0.5, 0.7),
                     (1e15, 0.75, 0.9)]:
        p = O.default_params(hmf_amp=s8)
        assert L.y3o_hmf_formula(C.byref(p), math.log(M), z) == pytest.approx(hand(M, z, s8),
                                                                              rel=1e-13)
    # the table the integrand interpolates holds the formula at its knots
    p = O.default_params(hmf_amp=0.75)
    t = O.OracleTables(p, O.spec_wide())
    for i, j in [(0, 0), (57, 11), (199, 31)]:
        M, z = math.exp(t.lnm[i]), t.z[j]
        assert t.hmf[j * len(t.lnm) + i] == pytest.approx(hand(M, z, 0.75), rel=1e-13)


def test_mor_lognormal_moments():
    """<ln lambda> = 3.2 + slope ln(M / 3e14) + 0.5 ln((1+z)/1.35) and
    Var(ln lambda) = sigma_ln^2, both by quadrature of the pdf in lambda."""
    for slope, sln, M, z in [(0.9, 0.25, 3e14, 0.35), (0.7, 0.1, 1e14, 0.2),
                             (1.1, 0.4, 2e15, 0.6), (0.9, 0.25, 5e13, 0.0)]:
        p = O.default_params(mor_slope=slope, mor_sigma_ln=sln)
        lnM = math.log(M)
        expect = 3.2 + slope * math.log(M / 3e14) + 0.5 * math.log((1 + z) / 1.35)
        f = lambda lt: L.y3o_mor_lognormal(C.byref(p), lt, lnM, z)
        c = math.exp(expect)
        pts = [c * math.exp(k * sln) for k in (-3, 0, 3)]
        m0 = integrate.quad(f, 0, np.inf, epsabs=0, epsrel=1e-12, limit=400, points=None)[0]
        lo, hi = c * math.exp(-14 * sln), c * math.exp(14 * sln)
        m1 = integrate.quad(lambda x: math.log(x) * f(x), lo, hi, epsabs=0, epsrel=1e-12,
                            limit=400, points=pts)[0]
        m2 = integrate.quad(lambda x: (math.log(x) - expect) ** 2 * f(x), lo, hi, epsabs=0,
                            epsrel=1e-12, limit=400, points=pts)[0]
        assert m0 == pytest.approx(1.0, abs=1e-9)
        assert m1 == pytest.approx(expect, abs=1e-9)
        assert m2 == pytest.approx(sln * sln, rel=1e-8)


def test_nfw_concentration_values():
    """c(M) = 5 (M / 1e14)^-0.1."""
    assert L.y3o_nfw_conc(math.log(1e14)) == pytest.approx(5.0, rel=1e-15)
    assert L.y3o_nfw_conc(math.log(1e15)) == pytest.approx(5.0 * 10 ** -0.1, rel=1e-14)
    assert L.y3o_nfw_conc(math.log(1e13)) == pytest.approx(5.0 * 10 ** 0.1, rel=1e-14)
    assert L.y3o_nfw_conc(math.log(3.7e14)) == pytest.approx(5.0 * 3.7 ** -0.1, rel=1e-14)


def test_zobs_pdf_width_and_bin():
    """z_obs ~ N(z_true, 0.01 (1 + z_true)): unit mass, mean z_true and
    standard deviation 0.01 (1 + z_true) by quadrature; the bin form is the
    normal CDF difference."""
    from scipy.stats import norm
    p = O.default_params()
    for zt in (0.2, 0.35, 0.6, 0.9):
        s = 0.01 * (1 + zt)
        f = lambda zo: L.y3o_zo_pdf(C.byref(p), zo, zt)
        a, b = zt - 12 * s, zt + 12 * s
        m0 = integrate.quad(f, a, b, epsabs=0, epsrel=1e-13, limit=200)[0]
        m1 = integrate.quad(lambda x: x * f(x), a, b, epsabs=0, epsrel=1e-13, limit=200)[0]
        m2 = integrate.quad(lambda x: (x - zt) ** 2 * f(x), a, b, epsabs=0, epsrel=1e-13,
                            limit=200)[0]
        assert m0 == pytest.approx(1.0, abs=1e-12)
        assert m1 == pytest.approx(zt, abs=1e-12)
        assert math.sqrt(m2) == pytest.approx(s, rel=1e-10)
        for lo, hi in [(0.2, 0.35), (0.35, 0.5), (0.5, 0.65)]:
            if lo >= zt:  # both ends above the mean: survival functions keep the tail
                ref = norm.sf(lo, zt, s) - norm.sf(hi, zt, s)
            else:
                ref = norm.cdf(hi, zt, s) - norm.cdf(lo, zt, s)
            assert L.y3o_zo_bin(C.byref(p), lo, hi, zt) == pytest.approx(ref, rel=1e-11,
                                                                          abs=1e-300)


def _chi_quad(om, z):
    """Flat LCDM comoving distance in Mpc/h by adaptive quadrature."""
    return 2997.92458 * integrate.quad(lambda x: 1 / math.sqrt(om * (1 + x) ** 3 + 1 - om),
                                       0, z, epsabs=0, epsrel=1e-13)[0]


def test_sigma_crit_inverse_table():
    """Sigma_crit^-1(z_l) from angular-diameter distances, the reference's
    form (src/modules/sigma_park_y1/buildSigmaCritInv.py:52-73): 4 pi G / c^2
    from its G = 4.51710305e-48 Mpc^3/Msun/s^2 and c = 9.71561189e-15 Mpc/s,
    D_ls = D_s - (1 + z_l)/(1 + z_s) D_l, Sigma_crit^-1 = 4 pi G/c^2 D_l D_ls / D_s;
    the table holds it in comoving units (the Delta-Sigma profile is comoving),
    i.e. times (1 + z_l)^2, in pc^2 / Msun with distances in Mpc/h; zero for
    z_l >= z_s.  D_A = chi / (1 + z) with chi by adaptive quadrature."""
    from scipy import constants as K
    G, c = 4.51710305e-48, 9.71561189e-15
    four_pi_g_c2 = 4 * math.pi * G / c ** 2 * 1e12  # Mpc/Msun -> pc^2/Msun per Mpc
    # the reference's constants agree with the IAU nominal solar GM (2015 B3)
    iau = 4 * math.pi * 1.3271244e20 / K.c ** 2 / K.parsec * 1e6  # pc/Msun * pc/Mpc
    assert four_pi_g_c2 == pytest.approx(iau, rel=2e-5)
    for om, zs in [(0.3, 1.2), (0.22, 1.2), (0.38, 0.6)]:
        p = O.default_params(omega_m=om, z_source=zs)
        t = O.OracleTables(p, O.spec_wide())
        ds = _chi_quad(om, zs) / (1 + zs)
        for i in (0, 1, 17, 64, 100, 127):
            zl = t.sci_z[i]
            dl = _chi_quad(om, zl) / (1 + zl)
            dls = ds - (1 + zl) / (1 + zs) * dl
            ref = four_pi_g_c2 * dl * dls / ds * (1 + zl) ** 2 if zl < zs else 0.0
            assert t.sci[i] == pytest.approx(ref, rel=1e-7, abs=1e-300), (om, zs, zl)
        # the lookup is the linear interpolation of the table, clamped
        zmid = 0.5 * (t.sci_z[40] + t.sci_z[41])
        assert L.y3o_sci(C.byref(t.t), zmid) == pytest.approx(0.5 * (t.sci[40] + t.sci[41]),
                                                              rel=1e-14)
        assert L.y3o_sci(C.byref(t.t), 5.0) == t.sci[-1]
    # sources in front of every lens knot above z_s: exactly zero
    p = O.default_params(z_source=0.5)
    t = O.OracleTables(p, O.spec_wide())
    assert np.all(t.sci[t.sci_z >= 0.5] == 0.0)
    assert np.all(t.sci[(t.sci_z > 0) & (t.sci_z < 0.5)] > 0.0)
