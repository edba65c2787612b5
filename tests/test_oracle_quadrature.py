"""Pins of the oracle's quadrature and integrators: polynomial exactness,
Random123 known-answer vectors, the reference's integrator closed forms
(test/integration.test.cc, ExampleScalarIntegrand.test.cc), and agreement of
the golden integrator with itself (panel orders) and with brute force."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O

L = O.lib


def test_gk21_polynomial_exactness():
    x = np.array(O.lib.y3o_gk21_nodes()[:11])
    wk = np.array(O.lib.y3o_gk21_wk()[:11])
    wg = np.array(O.lib.y3o_gk21_wg()[:5])
    nodes = np.concatenate([-x[:10], x[::-1]])
    wks = np.concatenate([wk[:10], wk[::-1]])
    gx = np.concatenate([-x[1:10:2], x[9::-2]])
    gws = np.concatenate([wg, wg[::-1]])
    for k in range(0, 32):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(wks, nodes ** k) - exact) < 1e-14
        if k < 20:
            assert abs(np.dot(gws, gx ** k) - exact) < 1e-14


def test_gl_rule_exactness():
    for n in (4, 6, 8):
        x = np.zeros(n)
        w = np.zeros(n)
        L.y3o_gl_rule(n, O.ptr(x), O.ptr(w))
        for k in range(2 * n):
            exact = 0.0 if k % 2 else 2.0 / (k + 1)
            assert abs(np.dot(w, x ** k) - exact) < 1e-14


def test_gk21_adaptive_sf1():
    # integration.test.cc sf1 = sin^2(pi x) on [0,1] = 0.5
    f = O._FN1(lambda x, c: math.sin(math.pi * x) ** 2)
    err = C.c_double()
    v = L.y3o_gk21_adapt(f, None, 0.0, 1.0, 1e-12, 0.0, C.byref(err))
    assert abs(v - 0.5) < 1e-14


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L.y3o_philox(c, k, o)
    return list(o)


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 10 rounds
    assert philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


@pytest.mark.parametrize("d", [2, 3, 4, 5, 6])
def test_genz_malik_exactness(d):
    rng = np.random.default_rng(d)
    lo = rng.uniform(-1, 0, d)
    hi = lo + rng.uniform(0.5, 2, d)
    c = np.ascontiguousarray(0.5 * (lo + hi))
    h = np.ascontiguousarray(0.5 * (hi - lo))
    for trial in range(12):
        # random monomial of total degree <= 7
        deg = rng.integers(0, 8)
        pw = np.zeros(d, dtype=int)
        for _ in range(deg):
            pw[rng.integers(0, d)] += 1
        f = O._FNN(lambda x, ctx: float(np.prod([x[i] ** pw[i] for i in range(d)])))
        exact = np.prod([(hi[i] ** (pw[i] + 1) - lo[i] ** (pw[i] + 1)) / (pw[i] + 1) for i in range(d)])
        r5 = C.c_double()
        four = np.zeros(d)
        r7 = L.y3o_gm_rule(f, None, d, O.ptr(c), O.ptr(h), C.byref(r5), O.ptr(four))
        assert abs(r7 - exact) < 1e-12 * max(1, abs(exact))
        if pw.sum() <= 5:
            assert abs(r5.value - exact) < 1e-12 * max(1, abs(exact))
    assert L.y3o_gm_num_nodes(d) == 1 + 4 * d + 2 * d * (d - 1) + 2 ** d


def _cub(which, d, lows, highs, eps=1e-3):
    f = O._FNN(lambda x, ctx: L.y3o_test_fn(which, x, d))
    r = O.Result()
    lo = np.ascontiguousarray(lows, dtype=float)
    hi = np.ascontiguousarray(highs, dtype=float)
    L.y3o_cubature_fn(f, None, d, O.ptr(lo), O.ptr(hi),
                      C.byref(O.cub_opts(eps_rel=eps, max_regions=20000, max_eval=2_000_000)),
                      C.byref(r))
    return r


def _veg(which, d, lows, highs, eps=1e-3):
    # VEGAS reports a 3-sigma error; 2e-3 keeps the sign-changing sf2_a_b
    # case within a CPU-test budget
    eps = max(eps, 2e-3)
    f = O._FNN(lambda x, ctx: L.y3o_test_fn(which, x, d))
    r = O.Result()
    lo = np.ascontiguousarray(lows, dtype=float)
    hi = np.ascontiguousarray(highs, dtype=float)
    L.y3o_vegas_fn(f, None, d, O.ptr(lo), O.ptr(hi),
                   C.byref(O.vegas_opts(eps_rel=eps, calls_per_iter=200000, itmax=30,
                                        max_eval=6_000_000)), C.byref(r))
    return r


@pytest.mark.parametrize("integrator", [_cub, _veg])
def test_integrator_closed_forms(integrator):
    # integration.test.cc:100-250 with epsrel 1e-3; ExampleScalarIntegrand.test.cc
    # (function, d, box, value quoted by the reference test, exact value)
    cases = [(1, 2, [0, 0], [1, 1], 1.0, 1.0),
             (1, 2, [-4, -11], [7, -3], -1276.0, -1276.0),
             (3, 2, [0, 0], [1, 1], 1.0, 1.0),
             (4, 2, [0, 0], [1, 1], 0.345833, 83.0 / 240.0),
             (4, 2, [-3, 4], [2, 7], 346.688, 346.6875)]
    for which, d, lo, hi, quoted, exact in cases:
        r = integrator(which, d, lo, hi)
        assert r.status == 0
        assert abs(r.value / quoted - 1) < 1e-3, (which, r.value, quoted)
        assert abs(r.value - exact) <= max(r.error, 1e-13 * abs(exact)), (which, r.value, r.error)


def test_cubature_sf6():
    # sf6 = 4.14199 (Mathematica, 6 figures), integration.test.cc:85-91
    r = _cub(2, 6, [1e-12] * 6, [1] * 6, eps=1e-4)
    assert r.status == 0
    assert abs(r.value - 4.14199) < 5e-5 * 4.14199 + 5e-6


def test_golden_panel_convergence():
    p = O.default_params(f_prj=0.0)
    t = O.OracleTables(p, O.spec_config0())
    vals = []
    for orders in [(3, 5), (4, 6), (6, 8)]:
        L.y3o_golden_set_orders(*orders)
        vals.append(O.golden(p, t, O.COUNT3D, [20., 30.])[0][0, 0])
    L.y3o_golden_set_orders(4, 6)
    assert max(vals) / min(vals) - 1 < 1e-11


def test_golden_vs_brute_force_count3d():
    p = O.default_params(f_prj=0.0)
    s = O.spec_config0()
    t = O.OracleTables(p, s)
    gold = O.golden(p, t, O.COUNT3D, [20., 30.])[0][0, 0]
    ig = O.make_integral(p, s, O.COUNT3D)
    # brute force of the bilinear integrand itself (hmf_smooth=0)
    r = O.cubature(p, t, ig, O.cub_opts(eps_rel=1e-7, max_regions=400000, max_eval=10**9,
                                        hmf_smooth=0))
    assert abs(r.value - gold) <= r.error
    assert abs(r.value / gold - 1) < 3e-7


def test_golden_vs_brute_force_count5d():
    p = O.default_params()
    s = O.spec_wide()
    t = O.OracleTables(p, s)
    lo = [20, 30, 45, 60, 300]
    zo = [0.2, 0.35, 0.5, 0.65]
    gold = O.golden(p, t, O.COUNT5D, lo, zo)[0]
    for i, j in [(0, 0), (2, 1)]:
        ig = O.make_integral(p, s, O.COUNT5D, lo[i], lo[i + 1], zo[j], zo[j + 1])
        r = O.cubature(p, t, ig, O.cub_opts(eps_rel=1e-4, max_regions=200000, max_eval=10**9))
        assert abs(r.value - gold[i, j]) <= r.error
        assert abs(r.value / gold[i, j] - 1) < 1e-4


def test_golden_vs_brute_force_dsigma6d():
    p = O.default_params()
    s = O.spec_wide()
    t = O.OracleTables(p, s)
    lo = [20, 30, 45, 60, 300]
    zo = [0.2, 0.35, 0.5, 0.65]
    R = 1.0
    gold = O.golden(p, t, O.COUNT5D, lo, zo, [R])[1]
    ig = O.make_integral(p, s, O.DSIGMA6D, lo[0], lo[1], zo[0], zo[1], R)
    r = O.cubature(p, t, ig, O.cub_opts(eps_rel=1e-3, max_regions=200000, max_eval=10**9))
    assert abs(r.value - gold[0, 0, 0]) <= r.error
    assert abs(r.value / gold[0, 0, 0] - 1) < 1e-3


def test_mapped_integrand_equals_physical():
    # v -> t window -> lt: f_mapped(y) = f(x) * lt * sigma_ln * (t_hi - t_lo),
    # with MOR(lt) dlt = phi(t) dt; Delta-Sigma: rmis from the Rayleigh CDF u
    # and theta on [0, pi] (factor 2 / P(rmis))
    p = O.default_params()
    s = O.spec_wide()
    t = O.OracleTables(p, s)
    L3 = 33.334803590584751
    rng = np.random.default_rng(3)
    for kind, (iv, iz, im) in [(O.COUNT5D, (2, 3, 4)), (O.DSIGMA6D, (1, 2, 3))]:
        ig = O.make_integral(p, s, kind, 20, 30, 0.2, 0.35, 0.8)
        for _ in range(20):
            y = np.array([rng.uniform(ig.lows[k], ig.highs[k]) for k in range(ig.ndim)])
            lnlbar = p.mor_slope * (y[im] - L3) + p.mor_lnl0 + p.mor_zslope * math.log((1 + y[iz]) / 1.35)
            a0 = (math.log(1.0) - lnlbar) / p.mor_sigma_ln
            b0 = (math.log(300.0) - lnlbar) / p.mor_sigma_ln
            # the window's smoothed clip: softplus of sharpness 4, never
            # wider than the hard clip [max(-7, a0), min(7, b0)] and within
            # ln 2 / 4 of it
            sp = lambda u: u if u > 30 else math.log1p(math.exp(u))
            tlo = a0 + sp(4 * (-7.0 - a0)) / 4
            thi = b0 - sp(4 * (b0 - 7.0)) / 4
            assert max(-7.0, a0) <= tlo <= max(-7.0, a0) + math.log(2) / 4 + 1e-12
            assert min(7.0, b0) - math.log(2) / 4 - 1e-12 <= thi <= min(7.0, b0)
            tt = tlo + (thi - tlo) * y[iv]
            lt = math.exp(lnlbar + p.mor_sigma_ln * tt)
            x = y.copy()
            x[iv] = lt
            jac = lt * p.mor_sigma_ln * (thi - tlo)
            if kind == O.DSIGMA6D:
                sg = ig.sigma_mis
                rm = sg * math.sqrt(-2 * math.log1p(-y[4]))
                x[4] = rm
                jac *= 2.0 / (rm / sg ** 2 * math.exp(-0.5 * rm * rm / sg ** 2))
            if kind == O.COUNT5D:
                # z_obs through w = (zo - zt) / sz over the bin's window,
                # clipped like t: zo = zt + sz w, dzo = sz (w_hi - w_lo) dv
                sz = p.sigz0 * (1 + y[iz])
                wa, wb = (0.2 - y[iz]) / sz, (0.35 - y[iz]) / sz
                wlo = wa + sp(4 * (-7.0 - wa)) / 4
                whi = wb - sp(4 * (wb - 7.0)) / 4
                x[1] = y[iz] + sz * (wlo + (whi - wlo) * y[1])
                if whi > wlo:  # an empty window (zt far outside the bin) maps to 0
                    assert 0.2 <= y[iz] + sz * wlo and y[iz] + sz * whi <= 0.35 + 1e-12
                    jac *= sz * (whi - wlo)
                else:
                    jac = 0.0
            phys = O.integrand(p, t, ig, x)
            mapped = L.y3o_integrand_mapped(C.byref(p), C.byref(t.t), C.byref(ig),
                                            O.ptr(np.ascontiguousarray(y)))
            assert abs(mapped - phys * jac) <= 1e-12 * abs(mapped) + 1e-300


def test_budget_extension_equals_larger_budget():
    """The budget extension continues the same refinement: an integral whose
    call limit would cut a step with the error above ext_ratio x tol gets
    the limit doubled (up to ext_regions) and ends exactly where a run with
    its final limit from the start ends (regions, evaluations, value); with
    the error within that ratio, or the ratio out of reach, it stops at the
    base limit."""
    p = O.default_params(f_prj=0.0)
    s = O.spec_config0()
    t = O.OracleTables(p, s)
    ig = O.make_integral(p, s, O.COUNT3D)
    nn = 33
    base, ext = 1024, 4096
    plain = O.cubature(p, t, ig, O.cub_opts(eps_rel=1e-9, max_regions=ext,
                                            max_eval=2 * base * nn))
    big = O.cubature(p, t, ig, O.cub_opts(eps_rel=1e-9, max_regions=ext, max_eval=2 * ext * nn))
    extd = O.cubature(p, t, ig, O.cub_opts(eps_rel=1e-9, max_regions=ext,
                                           max_eval=2 * base * nn, ext_regions=ext,
                                           ext_ratio=2.0))
    assert plain.neval <= 2 * base * nn < extd.neval
    assert (extd.neval, extd.nregions, extd.status) == (big.neval, big.nregions, big.status)
    assert extd.value == big.value and extd.error == big.error
    # the limit doubles while the error stays above the ratio: ~1.2e4 x tol
    # near 1024 regions, ~3.2e3 x tol near 2048 (64 splits per step, so the
    # decisions fall just before each limit), so a ratio of 6e3 doubles once
    # and the run equals one with 2048 regions from the start
    mid = O.cubature(p, t, ig, O.cub_opts(eps_rel=1e-9, max_regions=2 * ext,
                                          max_eval=2 * base * nn, ext_regions=2 * ext,
                                          ext_ratio=6e3, max_split_per_step=64))
    two = O.cubature(p, t, ig, O.cub_opts(eps_rel=1e-9, max_regions=2 * ext,
                                          max_eval=4 * base * nn, max_split_per_step=64))
    assert mid.neval == 4 * base * nn
    assert (mid.neval, mid.nregions, mid.value) == (two.neval, two.nregions, two.value)
    # the error at the base limit is ~1e-5 relative (1e4 x tol): a ratio
    # above that keeps the base limit
    far = O.cubature(p, t, ig, O.cub_opts(eps_rel=1e-9, max_regions=ext,
                                          max_eval=2 * base * nn, ext_regions=ext,
                                          ext_ratio=1e9))
    assert (far.neval, far.value) == (plain.neval, plain.value)
    # an integral that converges within its base limit is untouched
    loose = O.cubature(p, t, ig, O.cub_opts(eps_rel=1e-3, max_regions=ext,
                                            max_eval=2 * base * nn, ext_regions=ext,
                                            ext_ratio=2.0))
    ref = O.cubature(p, t, ig, O.cub_opts(eps_rel=1e-3, max_regions=ext, max_eval=2 * base * nn))
    assert ref.status == 0 and ref.neval < 2 * base * nn
    assert loose.status == 0 and (loose.neval, loose.value) == (ref.neval, ref.value)
