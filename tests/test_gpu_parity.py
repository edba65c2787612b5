"""GPU parity: the HIP path (through the C ABI of include/y3g.h) against the
oracle, the reference's fixtures and the golden values.

Tolerances: FP64 functors 1e-12 relative to the oracle (1e-6 / 1e-3 for the
reference fixtures, as in its Catch2 tests); FP32 functors 2e-5 relative;
integrals within their requested eps_rel of the golden value with a
reported error that bounds the true error; FP64 integrals equal to the
oracle's CPU restatement of the same algorithm to 1e-11.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle as O
import y3_cluster_cpp_amd as y3
from conftest import GOLDEN, load_columns

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "golden_fiducial.json")) as f:
    G = json.load(f)
LO, ZO, RADII = G["lo_edges"], G["zo_edges"], G["radii"]


def fid(rt, wide=True):
    p = y3.default_params() if wide else y3.default_params(f_prj=0.0)
    s = y3.spec_wide() if wide else y3.spec_config0()
    rt.tables_build(p, s)
    return p, s


def test_device_info(runtime):
    info = runtime.device_info()
    assert info["name"].startswith("gfx950")
    assert info["n_cu"] > 0


@pytest.mark.parametrize("wide", [False, True])
def test_tables_match_oracle(runtime, wide):
    p, s = fid(runtime, wide)
    hmf, da, sci = runtime.tables_download(0, s)
    op = O.default_params() if wide else O.default_params(f_prj=0.0)
    ot = O.OracleTables(op, O.spec_wide() if wide else O.spec_config0())
    assert np.max(np.abs(hmf / ot.hmf - 1)) < 1e-12
    assert np.max(np.abs(da[1:] / ot.da[1:] - 1)) < 1e-13 and da[0] == 0.0
    m = ot.sci > 0
    assert np.max(np.abs(sci[m] / ot.sci[m] - 1)) < 1e-12


def _fixture_ok(v, ref, prec):
    """FP64: the reference tests' own tolerance (Approx epsilon 1e-6,
    margin 1e-12).  FP32: 1e-4 relative plus 1e-7 of the largest value (the
    FP32 argument rounding of far Gaussian tails)."""
    if prec == y3.FP64:
        return np.abs(v - ref) <= np.maximum(1e-12, 1e-6 * np.abs(ref))
    return np.abs(v - ref) <= 1e-4 * np.abs(ref) + 1e-7 * np.max(np.abs(ref))


@pytest.mark.parametrize("prec", [y3.FP64, y3.FP32SPLIT])
def test_reference_fixtures_on_device(runtime, prec):
    # roffset_t.test.cc, lo_lc_t.test.cc, zo_zt_des_t.test.cc,
    # int_zo_zt_des_t.test.cc evaluated by the device functors
    d = load_columns("test_Rmis.txt", 2)
    v = runtime.eval_functor(y3.F_ROFFSET_GAMMA, d[:, :1], [0.15], precision=prec)
    assert np.all(_fixture_ok(v, d[:, 1], prec))
    d = load_columns("test_lolc_t.txt", 2)
    v = runtime.eval_functor(y3.F_LO_LC, d[:, :3], [1.66, 0.26, 1.43, 1.0], precision=prec)
    assert np.all(_fixture_ok(v * d[:, 1], d[:, 3], prec))
    d = load_columns("test_zo_zt_des.txt", 2)
    v = runtime.eval_functor(y3.F_SIGMA_PHOTOZ, d[:, 1:2], precision=prec)
    assert np.all(_fixture_ok(v, d[:, 2], prec))
    v = runtime.eval_functor(y3.F_ZO_ZT_DES, d[:, :2], precision=prec)
    assert np.all(_fixture_ok(v, d[:, 3], prec))
    d = load_columns("test_int_zo_zt_des.txt", 2)
    v = runtime.eval_functor(y3.F_INT_ZO_ZT_DES, d[:, :3], precision=prec)
    assert np.all(_fixture_ok(v, d[:, 4], prec))


def test_dv_do_dz_fixture_on_device(runtime):
    # dv_do_dz_t.test.cc: the reference D_A table uploaded as the table of a
    # point with Omega_m = 0.3, h = 0.7 (uniform 500-point grid on [0, 1])
    z = load_columns("z_da_test.txt")[:, 0]
    da = load_columns("d_a_test.txt")[:, 0]
    assert np.allclose(np.diff(z), z[1] - z[0], rtol=1e-9)
    s = y3.spec_wide()
    s.n_da = len(z)
    s.da_zmax = float(z[-1])
    p = y3.default_params()
    runtime.tables_build(p, s)
    runtime.tables_upload(0, da=da)
    d = load_columns("test_dv_do_dz.txt")
    for prec in (y3.FP64, y3.FP32SPLIT):
        v = runtime.eval_functor(y3.F_DV_DO_DZ, d[1:, :1], precision=prec)
        assert np.all(np.abs(v - d[1:, 1]) <= 1e-3 * np.abs(d[1:, 1]) + 1e-12)


def _oracle_fid():
    op = O.default_params()
    return op, O.OracleTables(op, O.spec_wide())


@pytest.mark.parametrize("prec,tol", [(y3.FP64, 1e-12), (y3.FP32SPLIT, 2e-5)])
def test_functors_match_oracle(runtime, prec, tol):
    p, s = fid(runtime)
    op, ot = _oracle_fid()
    rng = np.random.default_rng(7)
    n = 400
    lnM = rng.uniform(s.lnm_lo, s.lnm_hi, n)
    zt = rng.uniform(0.1, 0.75, n)
    lt = np.exp(rng.uniform(0.5, 5.5, n))
    lo = rng.uniform(20, 300, n)
    cases = [
        (y3.F_HMF, np.c_[lnM, zt], lambda r: O.lib.y3o_hmf(C.byref(op), C.byref(ot.t), r[0], r[1])),
        (y3.F_MOR, np.c_[lt, lnM, zt], lambda r: O.lib.y3o_mor_lognormal(C.byref(op), *r)),
        (y3.F_DV_DO_DZ, np.c_[zt], lambda r: O.lib.y3o_dvdz(C.byref(op), C.byref(ot.t), r[0])),
        (y3.F_SCI, np.c_[zt], lambda r: O.lib.y3o_sci(C.byref(ot.t), r[0])),
        (y3.F_LO_LT_PDF, np.c_[lo, lt], lambda r: O.lib.y3o_lo_lt_pdf(C.byref(op), *r)),
        (y3.F_LO_LT_BIN, np.c_[np.full(n, 20.), np.full(n, 30.), lt],
         lambda r: O.lib.y3o_lo_lt_bin(C.byref(op), *r)),
    ]
    for which, x, ref in cases:
        v = runtime.eval_functor(which, x, precision=prec)
        r = np.array([ref(row) for row in x])
        top = np.max(np.abs(r))
        # relative tolerance on values above 1e-8 (FP64) / 1e-4 (FP32) of
        # the largest, absolute below (far tails)
        cut = (1e-8 if prec == y3.FP64 else 1e-4) * top
        big = np.abs(r) > cut
        rel = np.abs(v[big] / r[big] - 1)
        assert np.all(rel <= tol), (which, rel.max())
        assert np.all(np.abs(v[~big] - r[~big]) <= tol * cut + 1e-300), which


@pytest.mark.parametrize("prec,tol", [(y3.FP64, 1e-11), (y3.FP32SPLIT, 5e-5)])
def test_nfw_functors_match_oracle(runtime, prec, tol):
    rng = np.random.default_rng(11)
    n = 300
    rho = 0.3 * 2.77533742639e11
    R = np.exp(rng.uniform(np.log(0.05), np.log(40), n))
    lnM = rng.uniform(30, 35.6, n)
    rmis = rng.uniform(0, 2, n)
    th = rng.uniform(0, 2 * np.pi, n)
    v = runtime.eval_functor(y3.F_NFW_DSIGMA, np.c_[R, lnM, np.full(n, rho), np.zeros(n)], precision=prec)
    r = np.array([O.lib.y3o_nfw_dsigma(a, b, rho, 0.0) for a, b in zip(R, lnM)])
    assert np.max(np.abs(v / r - 1)) < tol
    v = runtime.eval_functor(y3.F_DSIGMA_OFFSET, np.c_[R, rmis, th, lnM, np.full(n, rho), np.zeros(n)],
                             precision=prec)
    r = np.array([O.lib.y3o_dsigma_offset(a, b, c, d, rho, 0.0) for a, b, c, d in zip(R, rmis, th, lnM)])
    scale = np.array([O.lib.y3o_nfw_dsigma(a, d, rho, 0.0) for a, d in zip(R, lnM)])
    assert np.max(np.abs(v - r) / np.maximum(np.abs(r), 1e-3 * np.abs(scale))) < 10 * tol


def test_fp32_nfw_table_wide_range(runtime):
    """The FP32 NFW shape (knot table) over x = R / r_s from ~1e-6 to ~5e4,
    against the oracle's closed form: within 5e-6 relative (table 7e-7, the
    halo amplitude's FP32 arithmetic the rest)."""
    n = 2000
    rho = 0.3 * 2.77533742639e11
    R = np.geomspace(1e-6, 1e4, n)
    lnM = np.full(n, 32.2)
    v = runtime.eval_functor(y3.F_NFW_DSIGMA, np.c_[R, lnM, np.full(n, rho), np.zeros(n)],
                             precision=y3.FP32SPLIT)
    r = np.array([O.lib.y3o_nfw_dsigma(a, b, rho, 0.0) for a, b in zip(R, lnM)])
    assert np.max(np.abs(v / r - 1)) < 5e-6


@pytest.mark.parametrize("kind", [y3.COUNT3D, y3.COUNT5D, y3.DSIGMA6D])
@pytest.mark.parametrize("prec,tol", [(y3.FP64, 1e-12), (y3.FP32SPLIT, 5e-5)])
def test_mapped_integrand_matches_oracle(runtime, kind, prec, tol):
    wide = kind != y3.COUNT3D
    p, s = fid(runtime, wide)
    op = O.default_params() if wide else O.default_params(f_prj=0.0)
    ot = O.OracleTables(op, O.spec_wide() if wide else O.spec_config0())
    ig = y3.make_integral(p, s, kind, 20., 30., 0.2, 0.35, 1.0)
    oig = O.make_integral(op, O.spec_wide() if wide else O.spec_config0(), kind, 20., 30., 0.2, 0.35, 1.0)
    rng = np.random.default_rng(kind)
    Y = np.array([[rng.uniform(ig.lows[k], ig.highs[k]) for k in range(ig.ndim)] for _ in range(500)])
    r = np.array([O.lib.y3o_integrand_mapped(C.byref(op), C.byref(ot.t), C.byref(oig),
                                             O.ptr(np.ascontiguousarray(y))) for y in Y])
    m = np.abs(r) > 1e-6 * np.max(np.abs(r))
    # the direct form, the cubature's (zt, lnM)-factorised form and its FP32
    # node form (float coordinates)
    for which in (y3.F_INTEGRAND, y3.F_INTEGRAND_PAIRED, y3.F_INTEGRAND_NODE):
        v = runtime.eval_functor(which, Y, precision=prec, aux=ig)
        assert np.max(np.abs(v[m] / r[m] - 1)) < tol, which
        assert np.max(np.abs(v - r)) < 1e-4 * np.max(np.abs(r)), which


def _check(res, gold, eps):
    gold = np.asarray(gold)
    rel = np.abs(res["value"] - gold) / np.abs(gold)
    assert np.all(res["status"] == 0), res["status"]
    assert np.all(rel <= eps), rel.max()
    assert np.all(res["error"] >= np.abs(res["value"] - gold)), "error estimate below true error"


@pytest.mark.parametrize("prec", [y3.FP64, y3.FP32SPLIT])
@pytest.mark.parametrize("algo", [y3.PAGANI, y3.MCUBES])
@pytest.mark.parametrize("eps", [1e-3, 1e-4])
def test_config0_integral(runtime, prec, algo, eps):
    p, s = fid(runtime, wide=False)
    ig = y3.make_integral(p, s, y3.COUNT3D, 20., 30., eps_rel=eps, algorithm=algo)
    res = runtime.integrate([ig], y3.options(precision=prec, vegas_calls=6_000_000, vegas_itmax=5))
    _check(res, G["config0"]["value"], eps)


def test_config0_fp64_equals_cpu_restatement(runtime):
    p, s = fid(runtime, wide=False)
    op = O.default_params(f_prj=0.0)
    ot = O.OracleTables(op, O.spec_config0())
    oig = O.make_integral(op, O.spec_config0(), O.COUNT3D)
    for eps in (1e-3, 1e-5):
        ig = y3.make_integral(p, s, y3.COUNT3D, 20., 30., eps_rel=eps)
        g = runtime.integrate([ig], y3.options(precision=y3.FP64))[0]
        c = O.cubature(op, ot, oig, O.cub_opts(eps_rel=eps, max_regions=262144, max_eval=1 << 40))
        assert (g["nregions"], g["neval"]) == (c.nregions, c.neval), (eps, g, c.value, c.nregions)
        assert abs(g["value"] / c.value - 1) < 1e-11
    ig = y3.make_integral(p, s, y3.COUNT3D, 20., 30., eps_rel=1e-4, algorithm=y3.MCUBES)
    g = runtime.integrate([ig], y3.options(precision=y3.FP64, vegas_calls=200000, vegas_itmax=5))[0]
    c = O.vegas(op, ot, oig, O.vegas_opts(eps_rel=1e-4, calls_per_iter=200000, itmax=5,
                                          seed=0x59334331))
    assert g["neval"] == c.neval
    assert abs(g["value"] / c.value - 1) < 1e-11


@pytest.mark.parametrize("prec", [y3.FP64, y3.FP32SPLIT])
def test_config1_all_bins(runtime, prec):
    p, s = fid(runtime)
    igs = [y3.make_integral(p, s, y3.COUNT5D, LO[i], LO[i + 1], ZO[j], ZO[j + 1], eps_rel=1e-3)
           for i in range(4) for j in range(3)]
    res = runtime.integrate(igs, y3.options(precision=prec))
    _check(res, np.array(G["config1"]["values"]).reshape(-1), 1e-3)


def _check_within(res, gold, eps):
    """Every integral within eps of the golden value with an honest error;
    status 0 (converged) or 1 (region budget reached)."""
    gold = np.asarray(gold)
    rel = np.abs(res["value"] - gold) / np.abs(gold)
    assert set(np.unique(res["status"])) <= {0, 1}, res["status"]
    assert np.all(rel <= eps), (rel.max(), int(np.argmax(rel)))
    assert np.all(res["error"] >= np.abs(res["value"] - gold)), "error estimate below true error"


def test_config1_tight(runtime):
    """Config 1 at its own tolerance: all 12 five-dimensional bins at 1e-4
    (FP32 integrand, FP64 sums, the 32768-region budget)."""
    p, s = fid(runtime)
    gold = np.array(G["config1"]["values"]).reshape(-1)
    igs = [y3.make_integral(p, s, y3.COUNT5D, LO[i], LO[i + 1], ZO[j], ZO[j + 1], eps_rel=1e-4)
           for i in range(4) for j in range(3)]
    res = runtime.integrate(igs, y3.options(precision=y3.FP32SPLIT, max_regions=32768))
    _check_within(res, gold, 1e-4)


def test_config2_all_integrals(runtime):
    """Config 2 at its own tolerance: all 180 six-dimensional Delta-Sigma
    integrals (12 bins x 15 radii) at 1e-3, one launch."""
    p, s = fid(runtime)
    gold = np.array(G["config2"]["values"]).reshape(-1)
    igs = [y3.make_integral(p, s, y3.DSIGMA6D, LO[i], LO[i + 1], ZO[j], ZO[j + 1], RADII[k],
                            eps_rel=1e-3) for i in range(4) for j in range(3) for k in range(15)]
    res = runtime.integrate(igs, y3.options(precision=y3.FP32SPLIT, max_regions=32768))
    _check_within(res, gold, 1e-3)
    # most converge outright at this tolerance; those are within eps too
    assert np.mean(res["status"] == 0) > 0.5, np.mean(res["status"] == 0)


def test_results_independent_of_batch(runtime):
    p, s = fid(runtime)
    ig0 = y3.make_integral(p, s, y3.COUNT5D, LO[1], LO[2], ZO[1], ZO[2], eps_rel=1e-3)
    alone = runtime.integrate([ig0], y3.options(precision=y3.FP32SPLIT))[0]
    mix = [y3.make_integral(p, s, y3.COUNT5D, LO[i], LO[i + 1], ZO[j], ZO[j + 1],
                            eps_rel=1e-3 if (i + j) % 2 else 3e-4)
           for i in range(4) for j in range(3)] * 3
    mix[17] = ig0
    res = runtime.integrate(mix, y3.options(precision=y3.FP32SPLIT))
    assert res[17]["value"] == alone["value"] and res[17]["error"] == alone["error"]
    assert res[17]["nregions"] == alone["nregions"]


def test_error_paths(runtime):
    p, s = fid(runtime)
    ig = y3.make_integral(p, s, y3.COUNT5D, 20., 30., 0.2, 0.35)
    bad = y3.make_integral(p, s, y3.COUNT5D, 20., 30., 0.2, 0.35)
    bad.point = 5
    with pytest.raises(RuntimeError, match="malformed"):
        runtime.integrate([ig, bad])
    bad = y3.make_integral(p, s, y3.COUNT5D, 20., 30., 0.2, 0.35)
    bad.highs[0] = bad.lows[0]
    with pytest.raises(RuntimeError, match="empty volume"):
        runtime.integrate([bad])
    with pytest.raises(RuntimeError, match="options"):
        runtime.integrate([ig], y3.options(max_split_per_step=5000))
    # a tiny evaluation budget ends with status 1 and an honest error
    ig.max_eval = 20000
    r = runtime.integrate([ig])[0]
    assert r["status"] == 1 and r["neval"] <= 20000


@pytest.mark.parametrize("prec", [y3.FP64, y3.FP32SPLIT])
def test_plain_path_device_wide(runtime, prec):
    # one integral per launch sequence (the plain path), same tolerance bar
    p, s = fid(runtime)
    gold1 = np.array(G["config1"]["values"])
    gold2 = np.array(G["config2"]["values"])
    igs = [y3.make_integral(p, s, y3.COUNT5D, LO[0], LO[1], ZO[0], ZO[1], eps_rel=1e-4),
           y3.make_integral(p, s, y3.COUNT5D, LO[3], LO[4], ZO[2], ZO[3], eps_rel=1e-3),
           y3.make_integral(p, s, y3.DSIGMA6D, LO[1], LO[2], ZO[1], ZO[2], RADII[9], eps_rel=1e-3)]
    res = runtime.integrate(igs, y3.options(precision=prec, single=1))
    _check(res, np.array([gold1[0, 0], gold1[3, 2], gold2[1, 1, 9]]), np.array([1e-4, 1e-3, 1e-3]))
    ms, nev = runtime.last_timing()
    assert nev == res["neval"].sum()


def test_two_level_fp64_equals_cpu_restatement(runtime):
    """The optional two-level estimate follows the oracle's restatement:
    same regions, evaluations and value in FP64."""
    p, s = fid(runtime, wide=False)
    op = O.default_params(f_prj=0.0)
    ot = O.OracleTables(op, O.spec_config0())
    oig = O.make_integral(op, O.spec_config0(), O.COUNT3D)
    for eps in (1e-4, 1e-6):
        ig = y3.make_integral(p, s, y3.COUNT3D, 20., 30., eps_rel=eps)
        g = runtime.integrate([ig], y3.options(precision=y3.FP64, tl_c=2.0, tl_smin=0.125))[0]
        c = O.cubature(op, ot, oig, O.cub_opts(eps_rel=eps, max_regions=262144, max_eval=1 << 40,
                                               tl_c=2.0, tl_smin=0.125))
        assert (g["nregions"], g["neval"], g["status"]) == (c.nregions, c.neval, c.status)
        # the two-level estimate takes differences of nearly equal region sums,
        # which carry the last-bit differences of the integrand (the device
        # contracts the smoothed HMF's B-spline sums into FMAs): the errors
        # agree to ~1e-7 relative, the values to rounding
        assert abs(g["value"] / c.value - 1) < 1e-11 and abs(g["error"] / c.error - 1) < 1e-6


def test_large_batch_independence(runtime):
    """49,152 integrals (256 prior points) in one persistent launch: the
    first 16 points' results are bit-identical to a 16-point launch, every
    result is finite within the region budget, and the first 64 points'
    reported errors bound their true errors against the golden values."""
    from y3_cluster_cpp_amd import workloads as W
    with open(os.path.join(GOLDEN, "golden_points.json")) as f:
        gp = json.load(f)
    ref = np.concatenate([np.array(gp["counts"]), np.array(gp["dsigma"])], axis=1)
    pts = W.prior_points(256)
    assert np.allclose(pts[:64], np.array(gp["points"]))
    eps = np.minimum(np.vstack([W.eps_stream(256)]), 1e-4)
    spec = y3.spec_wide()
    params = [W.point_params(y3, pt) for pt in pts]
    opts = y3.options(max_regions=32768)

    def run(n):
        igs = []
        for k in range(n):
            igs += W.point_integrals(y3, params[k], spec, k, eps[k], y3.PAGANI)
        runtime.tables_build(params[:n], spec)
        return runtime.integrate((y3.Integral * len(igs))(*igs), opts)

    big = run(256)
    small = run(16)
    assert len(big) == 256 * 192
    for f in ("value", "error", "nregions", "neval", "status"):
        np.testing.assert_array_equal(big[f][:16 * 192], small[f], err_msg=f)
    assert np.all(np.isfinite(big["value"])) and np.all(np.isfinite(big["error"]))
    assert np.all(big["nregions"] <= 32768) and set(np.unique(big["status"])) <= {0, 1}
    head = big[:64 * 192]
    err = np.abs(head["value"] - ref.reshape(-1))
    assert np.all(head["error"] >= err)
    # tolerance on the stream: every converged integral (status 0) is within
    # its eps_rel; of all, at least 94% are within eps and counted (the
    # integrands read the smoothed HMF: median relative error ~3e-6 against
    # eps down to 1e-5; the bilinear table's knot aliasing held it at ~1e-5
    # and 88%)
    rel = err / np.abs(ref.reshape(-1))
    e = eps[:64].reshape(-1)
    conv = head["status"] == 0
    assert np.all(rel[conv] <= e[conv]), rel[conv].max()
    counted = (rel <= e) & (head["error"] >= err)
    assert counted.mean() >= 0.94, counted.mean()


def test_budget_extension(runtime):
    """Budget extension: an integral whose error is still above ext_ratio x
    its tolerance when its band budget is spent continues to ext_regions;
    one within that ratio stops at the band budget.  The batched FP64 path
    equals the oracle's CPU restatement (regions, evaluations, status,
    value), the plain path honours the same rule, and an ext_regions above
    max_regions is refused."""
    p, s = fid(runtime, wide=False)
    op = O.default_params(f_prj=0.0)
    ot = O.OracleTables(op, O.spec_config0())
    oig = O.make_integral(op, O.spec_config0(), O.COUNT3D)
    nn = 33
    base, ext = 1024, 4096
    ig = y3.make_integral(p, s, y3.COUNT3D, 20., 30., eps_rel=1e-9)
    # the error at the base budget is ~1.2e4 x the tolerance, ~3.2e3 x near
    # twice the budget: ratio 2 doubles the limit up to ext, 1e9 keeps the
    # base budget, 6e3 (64 splits per step) doubles it once
    for ratio, kmax, extended in ((2.0, 1024, True), (1e9, 1024, False), (6e3, 64, True)):
        o = dict(bands=[(1e-3, base)], max_regions=ext, ext_regions=ext, ext_ratio=ratio,
                 max_split_per_step=kmax)
        g = runtime.integrate([ig], y3.options(precision=y3.FP64, **o))[0]
        c = O.cubature(op, ot, oig, O.cub_opts(eps_rel=1e-9, max_regions=ext,
                                               max_eval=2 * base * nn, ext_regions=ext,
                                               ext_ratio=ratio, max_split_per_step=kmax))
        assert (g["nregions"], g["neval"], g["status"]) == (c.nregions, c.neval, c.status)
        assert abs(g["value"] / c.value - 1) < 1e-11
        if not extended:
            assert g["neval"] <= 2 * base * nn
        elif kmax == 64:
            assert g["neval"] == 4 * base * nn
        else:
            assert 2 * base * nn < g["neval"] <= 2 * ext * nn
        # the plain path and the FP32 integrand follow the same rule (their
        # sums differ from the FP64 batched ones in the last bits)
        for q in (dict(precision=y3.FP64, single=1), dict(precision=y3.FP32SPLIT)):
            h = runtime.integrate([ig], y3.options(**q, **o))[0]
            assert h["neval"] <= 2 * ext * nn
            assert (h["neval"] > 2 * base * nn) == extended
    with pytest.raises(RuntimeError, match="ext_regions"):
        runtime.integrate([ig], y3.options(max_regions=1024, ext_regions=2048, ext_ratio=2.0))


def test_tolerance_bands(runtime):
    """Tolerance-band budgets: an integral gets the budget of the first band
    whose edge is >= its eps_rel (per-integral call limit 2 x regions x
    nodes), the same integral outside every band the whole max_regions; the
    plain path honours the same limits; malformed bands are refused."""
    p, s = fid(runtime)
    igs = [y3.make_integral(p, s, y3.DSIGMA6D, LO[0], LO[1], ZO[0], ZO[1], RADII[5],
                            eps_rel=e) for e in (1e-9, 5e-5, 5e-4)]
    bands = [(1e-8, 8192), (1e-4, 2048)]
    r = runtime.integrate(igs, y3.options(bands=bands, max_regions=16384))
    nn = 149
    assert r["neval"][0] <= 2 * 8192 * nn and r["nregions"][0] > 4096
    # the call limit admits the initial regions (2 pieces on five axes, 4
    # along lnM) plus two per split: at most 2048 + n_init / 2 regions
    n_init = 2 ** 5 * 4
    assert r["neval"][1] <= 2 * 2048 * nn and r["nregions"][1] <= 2048 + n_init // 2
    assert r["nregions"][2] > r["nregions"][1] or r["status"][2] == 0  # outside: max_regions
    rp = runtime.integrate(igs[:2], y3.options(bands=bands, max_regions=16384, single=1,
                                               precision=y3.FP64))
    assert rp["neval"][0] <= 2 * 8192 * nn and rp["neval"][1] <= 2 * 2048 * nn
    with pytest.raises(RuntimeError, match="tolerance bands"):
        runtime.integrate(igs, y3.options(bands=[(1e-4, 2048), (1e-5, 4096)], max_regions=16384))
    with pytest.raises(RuntimeError, match="tolerance bands"):
        runtime.integrate(igs, y3.options(bands=[(1e-4, 32768)], max_regions=16384))


def test_vegas_dsigma6d_and_counts(runtime):
    """m-Cubes on the 6-D miscentred Delta-Sigma kind (one radius of each of
    config 2's 12 bins) and config 1's 12 counts, FP32 integrand, eps 1e-3:
    every reported error (3 sigma, inflated by chi2/dof) bounds the true
    error against the golden values, converged results meet their
    tolerance, and a second run repeats the values to the rounding of the
    FP64 histogram atomics."""
    p, s = fid(runtime)
    g1 = np.array(G["config1"]["values"]).reshape(-1)
    g2 = np.array(G["config2"]["values"])
    igs = [y3.make_integral(p, s, y3.COUNT5D, LO[i], LO[i + 1], ZO[j], ZO[j + 1], eps_rel=1e-3,
                            algorithm=y3.MCUBES) for i in range(4) for j in range(3)]
    igs += [y3.make_integral(p, s, y3.DSIGMA6D, LO[i], LO[i + 1], ZO[j], ZO[j + 1],
                             RADII[(3 * i + j) % 15], eps_rel=1e-3, algorithm=y3.MCUBES)
            for i in range(4) for j in range(3)]
    gold = np.concatenate([g1, [g2[i, j, (3 * i + j) % 15] for i in range(4) for j in range(3)]])
    o = y3.options(precision=y3.FP32SPLIT, vegas_calls=2_000_000, vegas_itmax=8)
    res = runtime.integrate(igs, o)
    err = np.abs(res["value"] - gold)
    assert np.all(res["error"] >= err), np.where(res["error"] < err)
    conv = res["status"] == 0
    assert np.all(err[conv] <= 1e-3 * np.abs(gold[conv])), (err / np.abs(gold))[conv].max()
    assert np.all(err <= 1e-2 * np.abs(gold)), (err / np.abs(gold)).max()
    again = runtime.integrate(igs, o)
    np.testing.assert_allclose(again["value"], res["value"], rtol=1e-8)


def test_tuned_stream_within_tolerance(runtime):
    """The benchmark's budgets (the committed tuning: bands, extension) on 16
    checked points of the stream (3,072 integrals, none of them the tuner's):
    at least 98.5% within their capped eps_rel with a reported error that
    bounds the true error (99.3% on the benchmark's 49,152).  The |R7 - R5|
    estimate is a heuristic: about 1 in 2,000 reported errors falls below
    the true error (16 of the benchmark's 49,152 at the base budgets), so
    at most 0.2% of these may, and at most 0.1% may report convergence
    outside their eps_rel."""
    from y3_cluster_cpp_amd import autotune
    from y3_cluster_cpp_amd import workloads as W
    with open(os.path.join(GOLDEN, "golden_stream.json")) as f:
        gs = json.load(f)
    idx = gs["index"][4:20]  # points 64 .. 304
    ref = np.array(gs["values"][4:20])
    pts = W.prior_points(4096)
    eps_all = np.minimum(W.eps_stream(4096), 1e-4)
    spec = y3.spec_wide()
    params = [W.point_params(y3, pts[k]) for k in idx]
    igs = []
    for local, k in enumerate(idx):
        igs += W.point_integrals(y3, params[local], spec, local, eps_all[k], y3.PAGANI)
    runtime.tables_build(params, spec)
    r = runtime.integrate((y3.Integral * len(igs))(*igs),
                          y3.options(precision=y3.FP32SPLIT, **autotune.load()))
    ref = ref.reshape(-1)
    eps = eps_all[idx].reshape(-1)
    err = np.abs(r["value"] - ref)
    rel = err / np.abs(ref)
    counted = (rel <= eps) & (r["error"] >= err)
    assert counted.mean() >= 0.985, counted.mean()
    conv = r["status"] == 0
    assert np.sum(rel[conv] > eps[conv]) <= 0.001 * len(igs), np.sort(rel[conv] / eps[conv])[-5:]
    assert np.sum(r["error"] < err) <= 0.002 * len(igs), np.sum(r["error"] < err)
