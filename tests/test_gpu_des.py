"""GPU: the device DES richness transfer (y3g_eval_functor
Y3G_F_INT_LC_LT_DES, INT_LC_LT_DES_t2) against the oracle on the
reference's tables, inside and outside the knots and on every lambda_c bin,
and the reference's own expected values; malformed tables are errors."""
import json
import os

import numpy as np
import pytest

import oracle as O
import y3_cluster_cpp_amd as y3
from conftest import REF_DATA
from test_oracle_des import lclt_par

pytestmark = pytest.mark.gpu


def test_int_lc_lt_des_device_matches_oracle(runtime):
    par, lt, zt = lclt_par()
    rng = np.random.default_rng(12)
    n = 4000
    x = np.column_stack([rng.uniform(5.0, 120.0, n), rng.uniform(-5.0, 220.0, n),
                         rng.uniform(-0.1, 0.9, n)])
    x[:40, 0] = [20, 30, 45, 60, 19.999, 29.999, 44.999, 59.999] * 5   # bin edges
    got = runtime.eval_functor(y3.F_INT_LC_LT_DES, x, par)
    ref = np.array([O.lib.y3o_int_lc_lt_des(O.ptr(par), *row) for row in x])
    assert np.all(np.abs(got - ref) <= 1e-14 + 1e-13 * np.abs(ref)), np.abs(got - ref).max()
    with open(os.path.join(REF_DATA, "int_lc_lt_des_answers.json")) as fh:
        a = json.load(fh)
    pts = np.array([(lc, l, z) for lc in a["lc"] for z in a["zt"] for l in a["lt"]])
    got = runtime.eval_functor(y3.F_INT_LC_LT_DES, pts, par)
    exp = np.array(a["answers"])
    assert np.all(np.abs(got - exp) <= 1e-5 + 1e-3 * np.abs(exp))


def test_int_lc_lt_des_rejects_bad_tables(runtime):
    par, _, _ = lclt_par()
    x = np.array([[25.0, 10.0, 0.3]])
    with pytest.raises(RuntimeError):
        runtime.eval_functor(y3.F_INT_LC_LT_DES, x, par[:-1])        # one value short
    bad = par.copy()
    bad[3] = bad[2]                                                  # lt knots not increasing
    with pytest.raises(RuntimeError):
        runtime.eval_functor(y3.F_INT_LC_LT_DES, x, bad)
    with pytest.raises(RuntimeError):
        runtime.eval_functor(y3.F_INT_LC_LT_DES, x[:, :2], par)      # two inputs


def test_mor_des_device_matches_oracle(runtime):
    """MOR_DES_t on the device against the oracle: masses below M_min
    (lambda = 1), lambdas beyond the tables' knots, two parameter sets."""
    rng = np.random.default_rng(31)
    n = 3000
    x = np.column_stack([rng.uniform(0.0, 400.0, n), rng.uniform(np.log(1e10), np.log(5e15), n),
                         rng.uniform(0.05, 0.9, n)])
    for q in (np.array([10 ** 11.2, 10 ** 12.4, 0.75, 0.2, 0.1, 0.3]),
              np.array([10 ** 11.5, 10 ** 12.9, 0.9, 1.5, -0.2, 0.45])):
        got = runtime.eval_functor(y3.F_MOR_DES, x, q)
        ref = np.array([O.lib.y3o_mor_des(O.ptr(q), *row) for row in x])
        assert np.all(np.abs(got - ref) <= 1e-290 + 1e-12 * np.abs(ref)), \
            np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300))
    with pytest.raises(RuntimeError):
        runtime.eval_functor(y3.F_MOR_DES, x[:4], q[:5])


def test_emg_des_device_matches_oracle(runtime):
    """EMG_DES_t pdf and cdf on the device against the oracle, across both
    branches of the reference's erfcx (|u| below and above 4), z beyond the
    curves' knots, ltr below 0.5, and +-inf for the cdf."""
    from test_oracle_des import emg_par
    par = emg_par()
    rng = np.random.default_rng(41)
    n = 3000
    ltr = rng.uniform(0.1, 150.0, n)
    x = np.column_stack([ltr * rng.uniform(0.0, 2.5, n), ltr, rng.uniform(0.0, 1.0, n)])
    x[:2, 0] = [np.inf, -np.inf]
    for which, fn in ((y3.F_EMG_DES_PDF, O.lib.y3o_emg_des_pdf),
                      (y3.F_EMG_DES_CDF, O.lib.y3o_emg_des_cdf)):
        rows = x if which == y3.F_EMG_DES_CDF else x[2:]
        got = runtime.eval_functor(which, rows, par)
        ref = np.array([fn(O.ptr(par), *row) for row in rows])
        assert np.all(np.abs(got - ref) <= 1e-290 + 1e-12 * np.abs(ref)), \
            (which, np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-290)))
    with pytest.raises(RuntimeError):
        runtime.eval_functor(y3.F_EMG_DES_CDF, x[:4], par[:-1])


def test_sig_max_device_matches_oracle(runtime):
    from test_oracle_des import sig_par
    par = sig_par()
    rng = np.random.default_rng(51)
    n = 3000
    x = np.column_stack([rng.uniform(0.01, 50.0, n), rng.uniform(29.0, 37.0, n),
                         rng.uniform(0.0, 1.2, n)])
    got = runtime.eval_functor(y3.F_SIG_MAX, x, par)
    ref = np.array([O.lib.y3o_sig_max(O.ptr(par), *row) for row in x])
    assert np.all(np.abs(got - ref) <= 1e-13 * np.maximum(1.0, np.abs(ref)))
    bad = par.copy()
    bad[4] = bad[3]                   # r axis not increasing
    with pytest.raises(RuntimeError):
        runtime.eval_functor(y3.F_SIG_MAX, x[:4], bad)
    with pytest.raises(RuntimeError):
        runtime.eval_functor(y3.F_SIG_MAX, x[:4], par[:-1])


# ------------------------------------------- the real-data Delta-Sigma integrand

def _des_setup(runtime):
    import ctypes as C
    from test_oracle_des import DES_BOX, des_data, des_integral
    p = y3.default_params()
    spec = y3.spec_wide()
    runtime.tables_build([p], spec)
    des = des_data()
    runtime.des_upload(0, des)
    lows = [b[0] for b in DES_BOX]
    highs = [b[1] for b in DES_BOX]
    ig = y3.make_integral_volume(p, spec, y3.DSIGMA6D_DES, lows, highs, (0.3, 0.45, 1.0),
                                 eps_rel=1e-3)
    op = O.default_params()
    ot = O.OracleTables(op, O.spec_wide())
    oig = des_integral(DES_BOX, zo=(0.3, 0.45), radius=1.0)
    return C, ig, des, op, ot, oig


def test_des_integrand_device_matches_oracle(runtime):
    C, ig, des, op, ot, oig = _des_setup(runtime)
    rng = np.random.default_rng(61)
    n = 2000
    from test_oracle_des import DES_BOX, des_axes
    x = np.column_stack([rng.uniform(l, h, n) for l, h in des_axes(DES_BOX)])
    got = runtime.eval_functor(y3.F_INTEGRAND, x, aux=ig)
    ref = np.array([O.lib.y3o_integrand_des(C.byref(op), C.byref(ot.t), O.ptr(des),
                                            C.byref(oig), O.ptr(np.ascontiguousarray(row)))
                    for row in x])
    scale = np.abs(ref).max()
    assert np.all(np.abs(got - ref) <= 1e-12 * np.abs(ref) + 1e-14 * scale), \
        np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300))
    with pytest.raises(RuntimeError):       # FP32 has no real-data path
        runtime.eval_functor(y3.F_INTEGRAND, x[:4], aux=ig, precision=y3.FP32SPLIT)


def test_des_cubature_matches_oracle(runtime):
    """The batched FP64 cubature and m-Cubes of the real-data integral
    against the oracle's cubature of the same integrand, within their
    errors; an FP32 request and a point without DES data are errors."""
    C, ig, des, op, ot, oig = _des_setup(runtime)
    r = runtime.integrate([ig], y3.options(precision=y3.FP64, max_regions=16384))
    o = O.Result()
    O.lib.y3o_cubature_des(C.byref(op), C.byref(ot.t), O.ptr(des), C.byref(oig),
                           C.byref(O.cub_opts(eps_rel=1e-4, max_regions=40_000,
                                              max_eval=2_000_000_000)), C.byref(o))
    v, e = r["value"][0], r["error"][0]
    assert abs(v - o.value) <= 2 * (e + o.error), (v, e, o.value, o.error)
    assert abs(v / o.value - 1) < 3e-3, (v, o.value)
    with pytest.raises(RuntimeError):
        runtime.integrate([ig], y3.options(precision=y3.FP32SPLIT))
    # m-Cubes (the reference's demo configuration runs this module with it)
    ig2 = y3.Integral.from_buffer_copy(ig)
    ig2.algorithm = y3.MCUBES
    ig2.max_eval = 20_000_000
    rv = runtime.integrate([ig2], y3.options(precision=y3.FP64, vegas_calls=2_000_000))
    v2, e2 = rv["value"][0], rv["error"][0]
    assert e2 > 0 and abs(v2 - o.value) <= 4 * e2 + 2 * o.error, (v2, e2, o.value, o.error)
    ig3 = y3.Integral.from_buffer_copy(ig)
    runtime.tables_build([y3.default_params(), y3.default_params(omega_m=0.28)], y3.spec_wide())
    ig3.point = 1
    with pytest.raises(RuntimeError):
        runtime.integrate([ig3], y3.options(precision=y3.FP64))


def test_des_module_end_to_end(runtime):
    """The reference's real-data Delta-Sigma module label through the
    CosmoSIS entry points, its keys read from the config and the sample:
    the result equals the library's on the same descriptor and the oracle's
    cubature within the errors."""
    import ctypes as C
    from y3_cluster_cpp_amd import cosmosis as cs
    from test_oracle_des import DES_BOX, SIG_M, SIG_R, SIG_Z, des_data, des_integral
    L = "SigmaMiscentY1MortCUDAScalarIntegrand"
    cfg = cs.DataBlock()
    for k, v in (("algorithm", "pagani"), ("eps_rel", 1e-3), ("eps_abs", 0.0),
                 ("max_eval", 100_000_000), ("use_cartesian_product", False),
                 ("max_regions", 16384)):
        cfg.put(L, k, v)
    for name, (lo, hi) in zip(("lo", "lc", "zt", "lnm", "rmis", "theta"), DES_BOX):
        cfg.put(L, name + "_low", [float(lo)])
        cfg.put(L, name + "_high", [float(hi)])
    cfg.put(L, "zo_low", [0.3])
    cfg.put(L, "zo_high", [0.45])
    cfg.put(L, "radii", [1.0])
    des = des_data()
    nr, nm, nz = len(SIG_R), len(SIG_M), len(SIG_Z)
    o = 11 + 3 + nr + nm + nz
    def make_sample():
        smp = cs.DataBlock()
        smp.put("cosmological_parameters", "omega_m", 0.3)
        for k, v in zip(("roffset_tau", "LO_LC_alpha", "LO_LC_a", "LO_LC_b", "LO_LC_R_lambda",
                         "mor_A", "mor_B", "mor_alpha", "mor_sigma", "mor_epsilon",
                         "z_mor_pivot"), des[:11]):
            smp.put("cluster_abundance", k, float(v))
        smp.put("deltasigma", "r_sigma_deltasigma", SIG_R)
        smp.put("deltasigma", "lnM", SIG_M)
        smp.put("matter_power_lin", "z", SIG_Z)
        smp.put("deltasigma", "sigma_1", des[o:o + nm * nr].reshape(nm, nr))
        smp.put("deltasigma", "sigma_2", des[o + nm * nr:o + nm * nr + nz * nr].reshape(nz, nr))
        smp.put("deltasigma", "bias", des[o + nm * nr + nz * nr:].reshape(nm, nz))
        return smp

    smp = make_sample()
    m = cs.Module(L)
    assert m.setup(cfg), "setup returned NULL"
    assert m.execute(smp) == 0
    vals, errs = smp.get_array(L, "vals"), smp.get_array(L, "errors")
    assert m.cleanup() == 0
    # the reference demo's settings for this module: algorithm = mcubes,
    # max_eval = 1e5 (y1_rerun/demo_sigma_mort_cuda.ini:139-142)
    cfg_mc = cs.DataBlock()
    for k, v in (("algorithm", "mcubes"), ("eps_rel", 1e-3), ("eps_abs", 0.0),
                 ("max_eval", 100_000), ("use_cartesian_product", False)):
        cfg_mc.put(L, k, v)
    for name, (lo, hi) in zip(("lo", "lc", "zt", "lnm", "rmis", "theta"), DES_BOX):
        cfg_mc.put(L, name + "_low", [float(lo)])
        cfg_mc.put(L, name + "_high", [float(hi)])
    cfg_mc.put(L, "zo_low", [0.3])
    cfg_mc.put(L, "zo_high", [0.45])
    cfg_mc.put(L, "radii", [1.0])
    mc = cs.Module(L)
    assert mc.setup(cfg_mc), "setup returned NULL"
    smp2 = make_sample()
    assert mc.execute(smp2) == 0
    vmc, emc = smp2.get_array(L, "vals"), smp2.get_array(L, "errors")
    assert mc.cleanup() == 0
    # the library on the same descriptor
    p = y3.default_params(omega_m=0.3)
    spec = y3.spec_wide()
    runtime.tables_build([p], spec)
    runtime.des_upload(0, des)
    ig = y3.make_integral_volume(p, spec, y3.DSIGMA6D_DES, [b[0] for b in DES_BOX],
                                 [b[1] for b in DES_BOX], (0.3, 0.45, 1.0), eps_rel=1e-3)
    ig.eps_abs, ig.max_eval = 0.0, 1 << 62
    r = runtime.integrate([ig], y3.options(precision=y3.FP64, max_regions=16384))
    assert vals.shape == (1,) and abs(vals[0] / r["value"][0] - 1) < 1e-12, (vals, r["value"])
    op = O.default_params(omega_m=0.3)
    ot = O.OracleTables(op, O.spec_wide())
    oo = O.Result()
    O.lib.y3o_cubature_des(C.byref(op), C.byref(ot.t), O.ptr(des),
                           C.byref(des_integral(DES_BOX)),
                           C.byref(O.cub_opts(eps_rel=1e-4, max_regions=40_000,
                                              max_eval=2_000_000_000)), C.byref(oo))
    assert abs(vals[0] - oo.value) <= 2 * (errs[0] + oo.error), (vals[0], errs[0], oo.value)
    assert emc[0] > 0 and abs(vmc[0] - oo.value) <= 5 * emc[0] + 2 * oo.error, \
        (vmc[0], emc[0], oo.value)


DES_MODULES = [  # label, kind, volume names, box, grid point
    ("SigmaCentY1MortCUDAScalarIntegrand", 4, ("lo", "zt", "lnm"), (1, 2, 3), (0.3, 0.45, 1.0)),
    ("NCCentY1MortCUDAScalarIntegrand", 5, ("lo", "zt", "lnm"), (1, 2, 3), (0.3, 0.45)),
    ("NCMiscentY1MortCUDAScalarIntegrand", 6, ("lo", "lc", "zt", "lnm", "rmis"),
     (0, 1, 2, 3, 4), (0.3, 0.45)),
]


@pytest.mark.parametrize("label,kind,names,axes,grid", DES_MODULES)
def test_des_other_modules(runtime, label, kind, names, axes, grid):
    """The other real-data modules (centred Delta-Sigma, centred and
    miscentred counts): the device integrand against the oracle at random
    points, and the module end to end against the library on the same
    descriptor and the oracle's cubature, within the errors."""
    import ctypes as C
    from y3_cluster_cpp_amd import cosmosis as cs
    from test_oracle_des import DES_BOX, SIG_M, SIG_R, SIG_Z, des_axes, des_data, des_integral
    box = [DES_BOX[a] for a in axes]
    des = des_data()
    nr, nm, nz = len(SIG_R), len(SIG_M), len(SIG_Z)
    o = 11 + 3 + nr + nm + nz
    sigma = kind == 4
    miscent = kind == 6
    up = des if sigma else des[:11]
    if not miscent:
        up = up.copy()
        up[:5] = 0.0
    p = y3.default_params(omega_m=0.3)
    spec = y3.spec_wide()
    runtime.tables_build([p], spec)
    runtime.des_upload(0, up)
    ig = y3.make_integral_volume(p, spec, kind, [b[0] for b in box], [b[1] for b in box], grid,
                                 eps_rel=1e-3)
    op = O.default_params(omega_m=0.3)
    ot = O.OracleTables(op, O.spec_wide())
    oig = des_integral(box, zo=grid[:2], radius=grid[2] if sigma else 1.0, kind=kind)
    rng = np.random.default_rng(70 + kind)
    x = np.column_stack([rng.uniform(l, h, 1500) for l, h in des_axes(box, kind)])
    got = runtime.eval_functor(y3.F_INTEGRAND, x, aux=ig)
    ref = np.array([O.lib.y3o_integrand_des(C.byref(op), C.byref(ot.t),
                                            O.ptr(np.ascontiguousarray(up)), C.byref(oig),
                                            O.ptr(np.ascontiguousarray(row))) for row in x])
    assert np.all(np.abs(got - ref) <= 1e-12 * np.abs(ref) + 1e-14 * np.abs(ref).max())

    cfg = cs.DataBlock()
    for k, v in (("algorithm", "pagani"), ("eps_rel", 1e-3), ("eps_abs", 0.0),
                 ("max_eval", 100_000_000), ("use_cartesian_product", False),
                 ("max_regions", 16384)):
        cfg.put(label, k, v)
    for name, (lo, hi) in zip(names, box):
        cfg.put(label, name + "_low", [float(lo)])
        cfg.put(label, name + "_high", [float(hi)])
    for name, v in zip(("zo_low", "zo_high", "radii"), grid):
        cfg.put(label, name, [float(v)])
    smp = cs.DataBlock()
    smp.put("cosmological_parameters", "omega_m", 0.3)
    keys = ["mor_A", "mor_B", "mor_alpha", "mor_sigma", "mor_epsilon", "z_mor_pivot"]
    for k, v in zip(keys, des[5:11]):
        smp.put("cluster_abundance", k, float(v))
    if miscent:
        for k, v in zip(("roffset_tau", "LO_LC_alpha", "LO_LC_a", "LO_LC_b", "LO_LC_R_lambda"),
                        des[:5]):
            smp.put("cluster_abundance", k, float(v))
    if sigma:
        smp.put("deltasigma", "r_sigma_deltasigma", SIG_R)
        smp.put("deltasigma", "lnM", SIG_M)
        smp.put("matter_power_lin", "z", SIG_Z)
        smp.put("deltasigma", "sigma_1", des[o:o + nm * nr].reshape(nm, nr))
        smp.put("deltasigma", "sigma_2", des[o + nm * nr:o + nm * nr + nz * nr].reshape(nz, nr))
        smp.put("deltasigma", "bias", des[o + nm * nr + nz * nr:].reshape(nm, nz))
    m = cs.Module(label)
    assert m.setup(cfg), "setup returned NULL"
    assert m.execute(smp) == 0
    vals, errs = smp.get_array(label, "vals"), smp.get_array(label, "errors")
    assert m.cleanup() == 0
    ig.eps_abs, ig.max_eval = 0.0, 1 << 62
    runtime.tables_build([p], spec)
    runtime.des_upload(0, up)
    r = runtime.integrate([ig], y3.options(precision=y3.FP64, max_regions=16384))
    assert abs(vals[0] / r["value"][0] - 1) < 1e-12, (vals, r["value"])
    oo = O.Result()
    O.lib.y3o_cubature_des(C.byref(op), C.byref(ot.t), O.ptr(np.ascontiguousarray(up)),
                           C.byref(oig), C.byref(O.cub_opts(eps_rel=1e-4, max_regions=40_000,
                                                            max_eval=2_000_000_000)),
                           C.byref(oo))
    assert abs(vals[0] - oo.value) <= 2 * (errs[0] + oo.error), (vals[0], errs[0], oo.value)


def test_card_module_and_integrand(runtime, monkeypatch):
    """numberCountsCard: MOR_DES_LOG_t and the card integrand on the device
    against the oracle, and the module (reading the reference's transfer
    tables from the working directory, as the reference does) end to end
    against the library and the oracle's cubature."""
    import ctypes as C
    from y3_cluster_cpp_amd import cosmosis as cs
    from test_oracle_des import DES_BOX, MOR_Q, card_data, des_integral, lclt_par
    ql = np.array([np.log10(MOR_Q[0]), np.log10(MOR_Q[1] / MOR_Q[0])] + list(MOR_Q[2:]))
    rng = np.random.default_rng(81)
    x = np.column_stack([rng.uniform(0.0, 300.0, 2000), rng.uniform(np.log(1e10), np.log(5e15), 2000),
                         rng.uniform(0.05, 0.9, 2000)])
    got = runtime.eval_functor(y3.F_MOR_DES_LOG, x, ql)
    ref = np.array([O.lib.y3o_mor_des_log(O.ptr(ql), *row) for row in x])
    assert np.all(np.abs(got - ref) <= 1e-290 + 1e-12 * np.abs(ref))

    lclt, _, _ = lclt_par()             # the reference's tables
    des = np.ascontiguousarray(np.concatenate([card_data()[:11], lclt]))
    box = [(np.log(20.0), np.log(60.0)), (np.log(5.0), np.log(150.0)), DES_BOX[2], DES_BOX[3]]
    p = y3.default_params(omega_m=0.3)
    spec = y3.spec_wide()
    runtime.tables_build([p], spec)
    runtime.des_upload(0, des)
    ig = y3.make_integral_volume(p, spec, y3.COUNT4D_CARD, [b[0] for b in box],
                                 [b[1] for b in box], (0.3, 0.45), eps_rel=1e-3)
    op = O.default_params(omega_m=0.3)
    ot = O.OracleTables(op, O.spec_wide())
    oig = des_integral(box, kind=7)
    xs = np.column_stack([rng.uniform(l, h, 1500) for l, h in box])
    got = runtime.eval_functor(y3.F_INTEGRAND, xs, aux=ig)
    ref = np.array([O.lib.y3o_integrand_des(C.byref(op), C.byref(ot.t), O.ptr(des), C.byref(oig),
                                            O.ptr(np.ascontiguousarray(row))) for row in xs])
    assert np.all(np.abs(got - ref) <= 1e-12 * np.abs(ref) + 1e-14 * np.abs(ref).max())

    L = "numberCountsCard"
    cfg = cs.DataBlock()
    for k, v in (("algorithm", "pagani"), ("eps_rel", 1e-3), ("eps_abs", 0.0),
                 ("max_eval", 100_000_000), ("use_cartesian_product", False),
                 ("max_regions", 16384)):
        cfg.put(L, k, v)
    for name, (lo, hi) in zip(("lnLo", "lnLt", "zt", "lnm"), box):
        cfg.put(L, name + "_low", [float(lo)])
        cfg.put(L, name + "_high", [float(hi)])
    cfg.put(L, "zo_low", [0.3])
    cfg.put(L, "zo_high", [0.45])
    smp = cs.DataBlock()
    smp.put("cosmological_parameters", "omega_m", 0.3)
    for k, v in zip(("mor_logMmin", "mor_logRatio", "mor_alpha", "mor_sigma", "mor_epsilon",
                     "z_mor_pivot"), ql):
        smp.put("cluster_abundance", k, float(v))
    monkeypatch.chdir(REF_DATA)
    m = cs.Module(L)
    assert m.setup(cfg), "setup returned NULL"
    assert m.execute(smp) == 0
    vals, errs = smp.get_array(L, "vals"), smp.get_array(L, "errors")
    assert m.cleanup() == 0
    des2 = np.ascontiguousarray(np.concatenate([[0.0] * 5, ql, lclt]))
    runtime.tables_build([p], spec)
    runtime.des_upload(0, des2)
    ig.eps_abs, ig.max_eval = 0.0, 1 << 62
    r = runtime.integrate([ig], y3.options(precision=y3.FP64, max_regions=16384))
    assert abs(vals[0] / r["value"][0] - 1) < 1e-12, (vals, r["value"])
    oo = O.Result()
    O.lib.y3o_cubature_des(C.byref(op), C.byref(ot.t), O.ptr(des2), C.byref(oig),
                           C.byref(O.cub_opts(eps_rel=1e-4, max_regions=40_000,
                                              max_eval=2_000_000_000)), C.byref(oo))
    assert abs(vals[0] - oo.value) <= 2 * (errs[0] + oo.error), (vals[0], errs[0], oo.value)


def test_des_data_discarded_by_table_rebuild(runtime):
    """DES data belongs to one sample: rebuilding the tables (same count and
    spec) or uploading a point's tables discards it, so a real-data
    integral can never read the previous sample's data."""
    from y3_cluster_cpp_amd.workloads import des_synthetic, des_volumes
    p = y3.default_params()
    spec = y3.spec_wide()
    runtime.tables_build([p], spec)
    runtime.des_upload(0, des_synthetic())
    lows, highs, grid = des_volumes()[0]
    ig = y3.make_integral_volume(p, spec, y3.DSIGMA6D_DES, lows, highs, grid, eps_rel=1e-2)
    o = y3.options(precision=y3.FP64, max_regions=256)
    assert np.isfinite(runtime.integrate([ig], o)["value"][0])
    runtime.tables_build([p], spec)
    with pytest.raises(RuntimeError, match="no DES data"):
        runtime.integrate([ig], o)
    runtime.des_upload(0, des_synthetic())
    runtime.integrate([ig], o)
    hmf, da, sci = runtime.tables_download(0, spec)
    runtime.tables_upload(0, hmf=hmf)
    with pytest.raises(RuntimeError, match="no DES data"):
        runtime.integrate([ig], o)


def test_des_cubature_real_table_sizes(runtime):
    """SIG_MAX tables at the reference's real sizes (1000 radii x 300 masses,
    data/deltasigma_r_perp.txt, deltasigma_m_h.txt; 10 redshifts): far past
    the workgroup's LDS copy (2048 doubles), so the kernel reads them from
    global memory.  The tables hold the same bilinear functions as the small
    ones (bilinear interpolation reproduces them), so the integral is the
    small tables' one; the device FP64 cubature follows the oracle's step for
    step at the same budget."""
    import ctypes as C
    from test_oracle_des import BI, DES_BOX, S1, S2, des_data, des_integral
    # the small tables' ranges (so that both clamp alike), 1000 x 300 x 10 knots
    r = np.logspace(np.log10(0.1), np.log10(30.0), 1000)
    m = np.linspace(30.0, 36.0, 300)
    z = np.linspace(0.1, 1.0, 10)

    def tab(xs, ys, c):
        X, Y = np.meshgrid(xs, ys)
        return (c[0] + c[1] * X + c[2] * Y + c[3] * X * Y).reshape(-1)
    head = des_data()[:11]
    big = np.ascontiguousarray(np.concatenate(
        [head, [len(r), len(m), len(z)], r, m, z, tab(r, m, S1), tab(r, z, S2), tab(z, m, BI)]))
    assert big.size > 300_000
    p = y3.default_params()
    spec = y3.spec_wide()
    runtime.tables_build([p], spec)
    runtime.des_upload(0, big)
    lows = [b[0] for b in DES_BOX]
    highs = [b[1] for b in DES_BOX]
    ig = y3.make_integral_volume(p, spec, y3.DSIGMA6D_DES, lows, highs, (0.3, 0.45, 1.0),
                                 eps_rel=1e-9)
    g = runtime.integrate([ig], y3.options(precision=y3.FP64, max_regions=2048))[0]
    op = O.default_params()
    ot = O.OracleTables(op, O.spec_wide())
    oig = des_integral(DES_BOX, zo=(0.3, 0.45), radius=1.0)
    o = O.Result()
    O.lib.y3o_cubature_des(C.byref(op), C.byref(ot.t), O.ptr(big), C.byref(oig),
                           C.byref(O.cub_opts(eps_rel=1e-9, max_regions=2048,
                                              max_eval=1 << 40)), C.byref(o))
    assert (g["nregions"], g["neval"]) == (o.nregions, o.neval)
    assert abs(g["value"] / o.value - 1) < 1e-10, (g["value"], o.value)
    # the small tables' integral, same budget
    runtime.des_upload(0, des_data())
    s = runtime.integrate([ig], y3.options(precision=y3.FP64, max_regions=2048))[0]
    assert abs(g["value"] / s["value"] - 1) < 1e-9, (g["value"], s["value"])
