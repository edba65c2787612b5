"""The oracle's DES richness transfer INT_LC_LT_DES_t2
(int_lc_lt_des_t2.cuh:48-63) against the reference's own test
(test/int_lc_lt_des_t_test.test.cu: 720 expected values at the table knots,
compared there with epsilon 1e-3 and margin 1e-5) and against what the
mathematics fixes: exactness on bilinear tables between knots, the lambda_c
bin edges, and clamping outside the knots."""
import ctypes as C
import json
import os

import numpy as np
import pytest

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


# ---------------------------------------------------------------- MOR_DES_t

# cluster_abundance/{mor_A, mor_B, mor_alpha, mor_sigma, mor_epsilon,
# z_mor_pivot}: a DES-like point (A = M_min, B = M_1 in M_sun/h)
MOR_Q = np.array([10 ** 11.2, 10 ** 12.4, 0.75, 0.2, 0.1, 0.3])


def mor_sigma_skew(si, lam):
    s, k = np.zeros(1), np.zeros(1)
    O.lib.y3o_mor_des_sigma_skew(si, lam, O.ptr(s), O.ptr(k))
    return s[0], k[0]


def test_mor_des_tables_read_by_hand():
    """Entries of sig_skewnorml_flat and skews at knots, read by hand from
    mor_des_t.cuh (row = sigma_intr knot, column = lambda_sat knot), and the
    bilinear value at a cell centre (the mean of its corners)."""
    cases = [(0.05, 0.5, 1.211319, 9.120214), (0.05, 244.0, 19.830625, 0.053913),
             (2.0, 244.0, 488.220129, 0.002402), (0.5, 0.7, 1.282591, 2.759347),
             (0.1, 14.0, 4.061805, 0.250297), (1.3, 24.0, 31.606019, 0.039318),
             (1.7, 24.0, 41.117558, 0.030344)]
    for si, lam, sig, skw in cases:
        s, k = mor_sigma_skew(si, lam)
        assert abs(s - sig) < 1e-12 and abs(k - skw) < 1e-12, (si, lam, s, k)
    s, _ = mor_sigma_skew(0.075, 0.55)
    assert abs(s - (1.211319 + 1.307980 + 1.211227 + 1.309620) / 4) < 1e-12
    # clamped beyond the knots
    assert mor_sigma_skew(5.0, 1e4) == mor_sigma_skew(2.0, 244.0)


def _mor(lt, lnM, zt, q=MOR_Q):
    return O.lib.y3o_mor_des(O.ptr(np.ascontiguousarray(q)), lt, lnM, zt)


def _lam(lnM, zt, q=MOR_Q):
    v = ((np.exp(lnM) - q[0]) / (q[1] - q[0])) ** q[2] * ((1 + zt) / (1 + q[5])) ** q[4]
    return max(1.0, v)


def test_mor_des_skew_normal_moments():
    """N(x; 0, s) erfc(-a x / (sqrt2 s)) is the skew-normal density: unit
    mass over lt, mean lambda + s d sqrt(2/pi) with d = a / sqrt(1 + a^2)."""
    from scipy import integrate
    for lnM, zt in ((np.log(3e13), 0.3), (np.log(2e14), 0.55), (np.log(1e15), 0.7),
                    (np.log(1e11), 0.4)):   # M < A: lambda = 1
        lam = _lam(lnM, zt) if np.exp(lnM) > MOR_Q[0] else 1.0
        s, a = mor_sigma_skew(MOR_Q[3], lam)
        lo, hi = lam - 40 * s, lam + 40 * s
        mass = integrate.quad(lambda t: _mor(t, lnM, zt), lo, hi, points=[lam], limit=200,
                              epsabs=0, epsrel=1e-11)[0]
        mean = integrate.quad(lambda t: t * _mor(t, lnM, zt), lo, hi, points=[lam], limit=200,
                              epsabs=0, epsrel=1e-11)[0]
        d = a / np.sqrt(1 + a * a)
        assert abs(mass - 1) < 1e-9, (lnM, zt, mass)
        assert abs(mean - (lam + s * d * np.sqrt(2 / np.pi))) < 1e-8 * max(1.0, lam), (lnM, mean)


# ---------------------------------------------------------------- EMG_DES_t

EMG_Z = np.array([0.1, 0.3, 0.5, 0.8])


def emg_par(curves=None):
    """par = {n, z, a_mu, b_mu, a_sig, b_sig, a_tau, b_tau, a_fprj, b_fprj}:
    DES-like curves (mu ~ l, sigma ~ l^0.5, tau ~ 1/l, fprj ~ 0.2) that vary
    linearly in z."""
    if curves is None:
        t = (EMG_Z - 0.1) / 0.7
        curves = [0.5 + 0.5 * t, 0.95 + 0.1 * t, 0.5 + 0.05 * t, 1.0 + 0.2 * t,
                  0.9 + 0.1 * t, 2.0 - 0.5 * t, 1.0 + 0 * t, 0.2 + 0.1 * t]
    return np.ascontiguousarray(np.concatenate([[len(EMG_Z)], EMG_Z] + list(curves)))


def emg_params(par, ltr, z):
    out = np.zeros(4)
    O.lib.y3o_emg_des_params(O.ptr(par), ltr, z, O.ptr(out))
    return out


def test_emg_des_curves_linear_between_knots_and_clamped():
    par = emg_par()
    for z in (0.1, 0.2, 0.45, 0.8):
        t = (z - 0.1) / 0.7
        mu, s, tau, f = emg_params(par, 10.0, z)
        assert abs(mu - ((0.5 + 0.5 * t) + (0.95 + 0.1 * t) * 10.0)) < 1e-12
        assert abs(s - (1.0 + 0.2 * t) * 10.0 ** (0.5 + 0.05 * t)) < 1e-12
        assert abs(tau - (2.0 - 0.5 * t) / 10.0 ** (0.9 + 0.1 * t)) < 1e-12
        assert abs(f - (0.2 + 0.1 * t) / (1 + np.exp(-10.0))) < 1e-12
    assert np.all(emg_params(par, 10.0, 1.5) == emg_params(par, 10.0, 0.8))
    assert np.all(emg_params(par, 0.1, 0.3) == emg_params(par, 0.5, 0.3))   # ltr >= 0.5


def test_emg_des_pdf_cdf_consistent_and_moments():
    """cdf' = pdf; unit mass; with the projection fraction forced to 0 and
    to 1 (b_fprj = 0 / 1e9, clamped to 1) the mean and variance are the
    Gaussian's and the EMG's (mu + 1/tau, sigma^2 + 1/tau^2)."""
    from scipy import integrate
    par = emg_par()
    for ltr, z in ((5.0, 0.2), (25.0, 0.45), (80.0, 0.7)):
        mu, s, tau, f = emg_params(par, ltr, z)
        pdf = lambda x: O.lib.y3o_emg_des_pdf(O.ptr(par), x, ltr, z)
        cdf = lambda x: O.lib.y3o_emg_des_cdf(O.ptr(par), x, ltr, z)
        lo, hi = mu - 30 * s, mu + 30 * s + 60 / tau
        assert abs(integrate.quad(pdf, lo, hi, points=[mu], limit=400, epsrel=1e-12)[0] - 1) < 1e-8
        for x in np.linspace(mu - 3 * s, mu + 3 * s + 2 / tau, 9):
            h = 1e-4 * s
            d = (cdf(x + h) - cdf(x - h)) / (2 * h)
            assert abs(d - pdf(x)) < 2e-4 * pdf(x) + 1e-9, (ltr, z, x, d, pdf(x))
        assert cdf(-np.inf) == 0.0 and cdf(np.inf) == 1.0
        for bf, m1, v in ((0.0, mu, s * s), (1e9, mu + 1 / tau, s * s + 1 / tau ** 2)):
            curves = list(par[1 + len(EMG_Z):].reshape(8, -1))
            curves[7] = np.full(len(EMG_Z), bf)
            pq = emg_par(curves)
            g = lambda x, k: x ** k * O.lib.y3o_emg_des_pdf(O.ptr(pq), x, ltr, z)
            mean = integrate.quad(g, lo, hi, args=(1,), points=[mu], limit=400, epsrel=1e-12)[0]
            m2 = integrate.quad(g, lo, hi, args=(2,), points=[mu], limit=400, epsrel=1e-12)[0]
            assert abs(mean - m1) < 1e-7 * max(1, abs(m1)), (bf, mean, m1)
            assert abs(m2 - mean ** 2 - v) < 1e-6 * v, (bf, m2 - mean ** 2, v)


# ------------------------------------------------------------------ SIG_MAX

SIG_R = np.array([0.1, 0.3, 1.0, 3.0, 10.0, 30.0])
SIG_M = np.array([30.0, 31.5, 33.0, 34.5, 36.0])
SIG_Z = np.array([0.1, 0.4, 0.7, 1.0])
S1 = (2.0, 0.5, -0.05, 0.01)      # sigma_1 = c + a r + d lnM + e r lnM
S2 = (1.0, 0.8, 2.0, -0.2)        # sigma_2 = c + a r + d z + e r z
BI = (-0.4, 0.5, 0.02, 0.001)     # bias    = c + a z + d lnM + e z lnM


def sig_par():
    """Bilinear tables, each x-fastest: sigma_1 over (r, lnM), sigma_2 over
    (r, z), bias over (z, lnM)."""
    def tab(xs, ys, c):
        X, Y = np.meshgrid(xs, ys)   # [ny, nx], x fastest
        return c[0] + c[1] * X + c[2] * Y + c[3] * X * Y
    body = [tab(SIG_R, SIG_M, S1), tab(SIG_R, SIG_Z, S2), tab(SIG_Z, SIG_M, BI)]
    return np.ascontiguousarray(np.concatenate(
        [[len(SIG_R), len(SIG_M), len(SIG_Z)], SIG_R, SIG_M, SIG_Z] + [b.reshape(-1) for b in body]))


def test_sig_max_bilinear_branches_and_clamp():
    par = sig_par()
    f = lambda r, m, z: O.lib.y3o_sig_max(O.ptr(par), r, m, z)
    bil = lambda c, x, y: c[0] + c[1] * x + c[2] * y + c[3] * x * y
    rng = np.random.default_rng(6)
    wins = [0, 0]
    for _ in range(400):
        r, m, z = rng.uniform(0.1, 30.0), rng.uniform(30.0, 36.0), rng.uniform(0.1, 1.0)
        a, b = bil(S1, r, m), bil(BI, z, m) * bil(S2, r, z)
        wins[int(b > a)] += 1
        assert abs(f(r, m, z) - max(a, b)) < 1e-11 * max(1.0, abs(a), abs(b))
    assert min(wins) > 20, wins   # both branches exercised
    assert f(100.0, 40.0, 2.0) == f(30.0, 36.0, 1.0)
    assert f(0.01, 20.0, 0.0) == f(0.1, 30.0, 0.1)


# ------------------------------------------- the real-data Delta-Sigma integrand

def des_data(sig_const=None):
    """{roffset_tau, LO_LC alpha a b R_lambda, MOR_DES A B alpha sigma eps z_p,
    SIG_MAX}: with sig_const the SIG_MAX tables are constant (sigma_1 =
    sig_const, sigma_2 = 0), else the bilinear tables of sig_par()."""
    head = [0.2, 3.0, 0.3, 1.0, 0.9, 1e11, 10 ** 12.4, 0.75, 0.2, 0.1, 0.3]
    sp = sig_par()
    if sig_const is not None:
        nr, nm, nz = len(SIG_R), len(SIG_M), len(SIG_Z)
        o = 3 + nr + nm + nz
        sp = sp.copy()
        sp[o:o + nm * nr] = sig_const
        sp[o + nm * nr:o + nm * nr + nz * nr] = 0.0
    return np.ascontiguousarray(np.concatenate([head, sp]))


def des_axes(box, kind=3, reduced=True):
    """The integrated axes of a real-data integral over the reference volume
    `box`: the miscentred kinds (3, 6) as the device integrates them
    (reduced: lambda_obs in closed form, lo dropped) or all of them."""
    return box[1:] if reduced and kind in (3, 6) else box


def des_integral(box, zo=(0.3, 0.45), radius=1.0, kind=3, reduced=True):
    """An oracle descriptor over the reference volume `box`; reduced
    miscentred kinds keep lo's range in lo_lo / lo_hi."""
    ig = O.Integral()
    axes = des_axes(box, kind, reduced)
    ig.kind, ig.ndim = kind, len(axes)
    if len(axes) < len(box):
        ig.lo_lo, ig.lo_hi = box[0]
    ig.zo_lo, ig.zo_hi, ig.radius = zo[0], zo[1], radius
    for a, (l, h) in enumerate(axes):
        ig.lows[a], ig.highs[a] = l, h
    return ig


DES_BOX = [(0.0, 600.0), (5.0, 150.0), (0.25, 0.5), (np.log(1.3e14), np.log(4e14)),
           (0.2, 1.5), (0.0, 2 * np.pi)]


@pytest.mark.parametrize("reduced", [False, True])
def test_des_integrand_reduces_to_the_count_integral(reduced):
    """With a constant SIG_MAX the real-data integral over a box that holds
    the lo mass factorises: lo_lc integrates to 1 over lo, the angle gives
    2 pi / (2 x 3.1415926535897), the Gamma(2) offset integrates in closed
    form over [r1, r2], and what is left is
    int dzt dlnM dV/dz hmf Omega int_zo_zt P_MOR(lc box | lnM, zt) sigma,
    here by Gauss-Legendre products of the separately pinned functors.  The
    oracle's cubature must agree within its error, of the reference's 6-D
    integrand and of the 5-D one with lambda_obs in closed form."""
    p = O.default_params()
    t = O.OracleTables(p, O.spec_wide())
    des = des_data(sig_const=2.5)
    ig = des_integral(DES_BOX, reduced=reduced)
    r = O.Result()
    O.lib.y3o_cubature_des(C.byref(p), C.byref(t.t), O.ptr(des), C.byref(ig),
                           C.byref(O.cub_opts(eps_rel=1e-4, max_regions=40_000,
                                              max_eval=2_000_000_000)), C.byref(r))
    tau, (r1, r2) = des[0], DES_BOX[4]
    rfrac = (1 + r1 / tau) * np.exp(-r1 / tau) - (1 + r2 / tau) * np.exp(-r2 / tau)
    q = np.ascontiguousarray(des[5:11])

    def gl(n, a, b):
        x, w = np.polynomial.legendre.leggauss(n)
        return 0.5 * (b - a) * x + 0.5 * (a + b), 0.5 * (b - a) * w

    xz, wz = gl(24, *DES_BOX[2])
    xm, wm = gl(24, *DES_BOX[3])
    xl, wl = gl(400, *DES_BOX[1])
    exp = 0.0
    for z, a in zip(xz, wz):
        zf = (O.lib.y3o_dvdz(C.byref(p), C.byref(t.t), z) * O.lib.y3o_omega_z_des(z) *
              O.lib.y3o_int_zo_zt_des(0.3, 0.45, z))
        for m, b in zip(xm, wm):
            mfrac = sum(w * O.lib.y3o_mor_des(O.ptr(q), lc, m, z) for lc, w in zip(xl, wl))
            exp += a * b * zf * O.lib.y3o_hmf(C.byref(p), C.byref(t.t), m, z) * mfrac
    exp *= 2.5 * rfrac * 2 * np.pi / (2 * 3.1415926535897)
    assert abs(r.value - exp) <= max(2 * r.error, 1e-3 * exp), (r.value, exp, r.error, r.status)


def test_des_integrand_offset_radius_geometry():
    """The offset radius R' of the ring at angle theta: with sigma_1 linear
    in r (sigma_2 = 0) the angular integral is 2 pi (c + d lnM) + (a + e lnM)
    4 (R + rmis) E(k), k^2 = 4 R rmis / (R + rmis)^2 (E the complete elliptic
    integral of the second kind), scaled by the integrand at theta = pi,
    where R' = |R - rmis|."""
    from scipy import integrate, special
    p = O.default_params()
    t = O.OracleTables(p, O.spec_wide())
    des = des_data()
    nr, nm, nz = len(SIG_R), len(SIG_M), len(SIG_Z)
    o = 11 + 3 + nr + nm + nz
    des[o + nm * nr:o + nm * nr + nz * nr] = 0.0          # sigma_2 = 0
    for R, rm, lnM in ((1.0, 0.5, 33.0), (3.0, 0.8, 32.2), (0.7, 0.6, 34.1)):
        ig = des_integral(DES_BOX, radius=R, reduced=False)
        x = np.array([45.0, 40.0, 0.33, lnM, rm, 0.0])

        def f(th):
            x[5] = th
            return O.lib.y3o_integrand_des(C.byref(p), C.byref(t.t), O.ptr(des), C.byref(ig),
                                           O.ptr(x))

        c, a, d, e = S1
        s1 = lambda r: c + a * r + d * lnM + e * r * lnM
        K = f(np.pi) / s1(abs(R - rm))
        k2 = 4 * R * rm / (R + rm) ** 2
        exp = K * (2 * np.pi * (c + d * lnM) + (a + e * lnM) * 4 * (R + rm) * special.ellipe(k2))
        got = integrate.quad(f, 0, 2 * np.pi, epsabs=0, epsrel=1e-12, limit=200)[0]
        assert abs(got / exp - 1) < 1e-10, (R, rm, got, exp)


def test_lo_lc_bin_is_the_lo_integral_of_lo_lc():
    """lo_lc integrated over a lambda_obs bin in closed form against
    adaptive quadrature of lo_lc itself, on both sides of the ridge
    lo = mu lc and at small offsets (narrow ridge); a bin holding the whole
    Gaussian has unit mass; R = 0 is the indicator of mu lc in the bin."""
    from scipy import integrate
    par = (3.0, 0.3, 1.0, 0.9)                      # alpha, a, b, R_lambda
    rng = np.random.default_rng(17)
    for _ in range(60):
        lc = rng.uniform(5.0, 150.0)
        R = 10 ** rng.uniform(-3.0, 0.3)
        lo_lo = rng.uniform(0.5, 1.2) * lc
        lo_hi = lo_lo + rng.uniform(0.02, 0.5) * lc
        x = R / par[3]
        mu, sg = np.exp(-x * x / par[0] ** 2), par[1] * np.arctan(par[2] * x)
        pts = [p_ for p_ in (mu * lc - 3 * sg * lc, mu * lc, mu * lc + 3 * sg * lc)
               if lo_lo < p_ < lo_hi]
        ref = integrate.quad(lambda lo: O.lib.y3o_lo_lc(lo, lc, R, *par), lo_lo, lo_hi,
                             points=pts or None, epsabs=1e-300, epsrel=1e-12, limit=400)[0]
        got = O.lib.y3o_lo_lc_bin(lo_lo, lo_hi, lc, R, *par)
        assert abs(got - ref) <= 1e-10 * max(ref, 1e-12) + 1e-15, (lc, R, lo_lo, lo_hi, got, ref)
    from scipy.stats import norm
    x = 0.5 / par[3]
    mu, sg = np.exp(-x * x / par[0] ** 2), par[1] * np.arctan(par[2] * x)
    assert O.lib.y3o_lo_lc_bin(0.0, 1e4, 40.0, 0.5, *par) == pytest.approx(
        norm.sf(0.0, mu, sg) - norm.sf(1e4 / 40.0, mu, sg), abs=1e-15)   # 1 - 7.8e-11
    assert O.lib.y3o_lo_lc_bin(35.0, 45.0, 40.0, 0.0, *par) == 1.0
    assert O.lib.y3o_lo_lc_bin(20.0, 30.0, 40.0, 0.0, *par) == 0.0


@pytest.mark.parametrize("kind", [3, 6])
def test_reduced_integrand_is_the_lo_integral(kind):
    """The miscentred real-data integrands as the device integrates them
    (lambda_obs in closed form, descriptor without lo) equal the reference
    integrand (with lo) integrated over the lo range by adaptive quadrature,
    at random points of the other axes."""
    from scipy import integrate
    p = O.default_params()
    t = O.OracleTables(p, O.spec_wide())
    des = des_data()
    box = [(20.0, 30.0)] + (DES_BOX[1:] if kind == 3 else DES_BOX[1:5])
    red = des_integral(box, radius=1.0, kind=kind)
    full = des_integral(box, radius=1.0, kind=kind, reduced=False)
    rng = np.random.default_rng(40 + kind)
    for _ in range(25):
        y = np.array([rng.uniform(l, h) for l, h in box[1:]])
        got = O.lib.y3o_integrand_des(C.byref(p), C.byref(t.t), O.ptr(des), C.byref(red),
                                      O.ptr(np.ascontiguousarray(y)))

        def f(lo):
            x = np.ascontiguousarray(np.concatenate([[lo], y]))
            return O.lib.y3o_integrand_des(C.byref(p), C.byref(t.t), O.ptr(des),
                                           C.byref(full), O.ptr(x))
        ref = integrate.quad(f, *box[0], epsabs=1e-300, epsrel=1e-12, limit=400)[0]
        assert abs(got - ref) <= 1e-9 * abs(ref) + 1e-300, (y, got, ref)


@pytest.mark.parametrize("kind", [4, 5, 6])
def test_des_cent_and_count_integrands_reduce(kind):
    """The other real-data integrands reduce the same way: the centred
    Delta-Sigma (4) and counts (5) over (lo, zt, lnM) to the 2-D count
    integral with the MOR mass inside the lo box (times the constant
    SIG_MAX for 4), the miscentred counts (6) over (lo, lc, zt, lnM, rmis)
    to it times the offset's closed-form mass (lo_lc integrating to 1)."""
    p = O.default_params()
    t = O.OracleTables(p, O.spec_wide())
    des = des_data(sig_const=2.5)
    box = {4: [DES_BOX[1], DES_BOX[2], DES_BOX[3]], 5: [DES_BOX[1], DES_BOX[2], DES_BOX[3]],
           6: DES_BOX[:5]}[kind]
    ig = des_integral(box, radius=1.0, kind=kind, reduced=False)
    r = O.Result()
    O.lib.y3o_cubature_des(C.byref(p), C.byref(t.t), O.ptr(des), C.byref(ig),
                           C.byref(O.cub_opts(eps_rel=1e-5, max_regions=40_000,
                                              max_eval=2_000_000_000)), C.byref(r))
    q = np.ascontiguousarray(des[5:11])
    lam_box = DES_BOX[1]

    def gl(n, a, b):
        x, w = np.polynomial.legendre.leggauss(n)
        return 0.5 * (b - a) * x + 0.5 * (a + b), 0.5 * (b - a) * w

    xz, wz = gl(24, *DES_BOX[2])
    xm, wm = gl(24, *DES_BOX[3])
    xl, wl = gl(400, *lam_box)
    exp = 0.0
    for z, a in zip(xz, wz):
        zf = (O.lib.y3o_dvdz(C.byref(p), C.byref(t.t), z) * O.lib.y3o_omega_z_des(z) *
              O.lib.y3o_int_zo_zt_des(0.3, 0.45, z))
        for m, b in zip(xm, wm):
            mfrac = sum(w * O.lib.y3o_mor_des(O.ptr(q), lc, m, z) for lc, w in zip(xl, wl))
            exp += a * b * zf * O.lib.y3o_hmf(C.byref(p), C.byref(t.t), m, z) * mfrac
    if kind == 4:
        exp *= 2.5
    if kind == 6:
        tau, (r1, r2) = des[0], DES_BOX[4]
        exp *= (1 + r1 / tau) * np.exp(-r1 / tau) - (1 + r2 / tau) * np.exp(-r2 / tau)
    assert abs(r.value - exp) <= max(2 * r.error, 2e-4 * exp), (kind, r.value, exp, r.error)


def test_mor_des_log_matches_mor_des_and_clamps():
    """MOR_DES_LOG_t with (log10 A, log10 (B/A)) is MOR_DES_t with (A, B)
    wherever M >= A and lambda >= 1; below M_min its lambda is 1 (the skew
    normal's mean there is 1 + sigma d sqrt(2/pi))."""
    from scipy import integrate
    qa = np.ascontiguousarray(MOR_Q)
    ql = np.ascontiguousarray([np.log10(MOR_Q[0]), np.log10(MOR_Q[1] / MOR_Q[0])] +
                              list(MOR_Q[2:]))
    rng = np.random.default_rng(9)
    for _ in range(200):
        lnM, zt = rng.uniform(np.log(3e13), np.log(2e15)), rng.uniform(0.1, 0.8)
        lt = rng.uniform(0.0, 300.0)
        a = O.lib.y3o_mor_des(O.ptr(qa), lt, lnM, zt)
        b = O.lib.y3o_mor_des_log(O.ptr(ql), lt, lnM, zt)
        assert abs(a - b) <= 1e-12 * abs(a) + 1e-300, (lnM, zt, lt, a, b)
    lnM, zt = np.log(1e11), 0.4   # below M_min
    s, k = mor_sigma_skew(MOR_Q[3], 1.0)
    f = lambda t: O.lib.y3o_mor_des_log(O.ptr(ql), t, lnM, zt)
    mass = integrate.quad(f, 1 - 40 * s, 1 + 40 * s, points=[1.0], epsrel=1e-11, limit=200)[0]
    mean = integrate.quad(lambda t: t * f(t), 1 - 40 * s, 1 + 40 * s, points=[1.0],
                          epsrel=1e-11, limit=200)[0]
    assert abs(mass - 1) < 1e-9
    assert abs(mean - (1 + s * k / np.sqrt(1 + k * k) * np.sqrt(2 / np.pi))) < 1e-8


def card_data(lclt_const=0.7):
    """DES data of the numberCountsCard integrand: MOR_DES_LOG parameters in
    the MOR places and INT_LC_LT tables (constant lclt_const, on the
    reference's knots) in place of SIG_MAX."""
    head = [0.0] * 5 + [11.2, 1.2, 0.75, 0.2, 0.1, 0.3]
    _, lt, zt = lclt_par()
    L, Z = np.meshgrid(lt, zt)
    par, _, _ = lclt_par([np.full_like(L, lclt_const)] * 4)
    return np.ascontiguousarray(np.concatenate([head, par]))


def test_card_integrand_reduces():
    """numberCountsCard over (ln lo, ln lt, zt, lnM) with constant transfer
    tables c: the ln lo integral is c times the box width, the ln lt one is
    int mor_log(lt) / lt dlt (no Jacobian is added: the variables are
    logarithms), leaving the 2-D count integral."""
    p = O.default_params()
    t = O.OracleTables(p, O.spec_wide())
    des = card_data(0.7)
    box = [(np.log(22.0), np.log(28.0)), (np.log(5.0), np.log(150.0)), DES_BOX[2], DES_BOX[3]]
    ig = des_integral(box, kind=7)
    r = O.Result()
    O.lib.y3o_cubature_des(C.byref(p), C.byref(t.t), O.ptr(des), C.byref(ig),
                           C.byref(O.cub_opts(eps_rel=1e-6, max_regions=40_000,
                                              max_eval=2_000_000_000)), C.byref(r))
    q = np.ascontiguousarray(des[5:11])

    def gl(n, a, b):
        x, w = np.polynomial.legendre.leggauss(n)
        return 0.5 * (b - a) * x + 0.5 * (a + b), 0.5 * (b - a) * w

    xz, wz = gl(24, *DES_BOX[2])
    xm, wm = gl(24, *DES_BOX[3])
    xl, wl = gl(400, *box[1])   # in ln lt
    exp = 0.0
    for z, a in zip(xz, wz):
        zf = (O.lib.y3o_dvdz(C.byref(p), C.byref(t.t), z) * O.lib.y3o_omega_z_des(z) *
              O.lib.y3o_int_zo_zt_des(0.3, 0.45, z))
        for m, b in zip(xm, wm):
            mm = sum(w * O.lib.y3o_mor_des_log(O.ptr(q), np.exp(u), m, z) for u, w in zip(xl, wl))
            exp += a * b * zf * O.lib.y3o_hmf(C.byref(p), C.byref(t.t), m, z) * mm
    exp *= 0.7 * (box[0][1] - box[0][0])
    # the 24 x 24 Gauss-Legendre reference is good to ~1e-4 over the
    # piecewise-bilinear HMF table
    assert abs(r.value - exp) <= max(2 * r.error, 1e-4 * exp), (r.value, exp, r.error)
