"""The Costanzi-2026 P[X] operator integrands on the device (Y3G_POP_*,
p_operator_gpu_t.cuh), through the C ABI and the CosmoSIS modules, against
the oracle (tests/test_oracle_pop.py pins it)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import y3_cluster_cpp_amd as y3
from test_oracle_pop import MOR, pop_data, pop_integral

pytestmark = pytest.mark.gpu
KINDS = (y3.POP_P1, y3.POP_I1, y3.POP_I2, y3.POP_J)
BOX = ([0.0, 0.1, 29.9], [130.0, 0.9, 35.7])


def _setup(runtime, des):
    p = y3.default_params()
    spec = y3.spec_wide()
    runtime.tables_build([p], spec)
    runtime.des_upload(0, des)
    op = O.default_params()
    ot = O.OracleTables(op, O.spec_wide())
    return p, spec, op, ot


def test_pop_integrand_matches_oracle(runtime):
    des = pop_data()
    p, spec, op, ot = _setup(runtime, des)
    rng = np.random.default_rng(9)
    n = 1500
    x = np.stack([rng.uniform(0.0, 60.0, n), rng.uniform(0.2, 0.6, n),
                  rng.uniform(30.0, 35.5, n)], 1)
    for kind in KINDS:
        for b, zo in [(0, (0.3, 0.45)), (3, (0.45, 0.6))]:
            ig = y3.make_integral_volume(p, spec, kind, *BOX, (zo[0], zo[1], float(b)))
            assert ig.lo_hi == [25.0, 37.5, 52.5, 130.0][b]
            got = runtime.eval_functor(y3.F_INTEGRAND, x, aux=ig)
            oig = pop_integral(kind, zo, b)
            ref = np.array([O.lib.y3o_integrand_des(C.byref(op), C.byref(ot.t), O.ptr(des),
                                                    C.byref(oig), O.ptr(np.ascontiguousarray(r)))
                            for r in x])
            assert np.count_nonzero(ref) > 100
            # 1e-8: the exclusion angle acos(c) has 1 - c ~ theta^2 / 2 ~ 1e-7
            # here, so the rounding of c (chi from a table lerp, FMA-contracted
            # on the device) moves the angle panel's edge by ~1e-9 relative
            scale = np.abs(ref).max()
            assert np.all(np.abs(got - ref) <= 1e-8 * np.abs(ref) + 1e-15 * scale), \
                (kind, b, np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)))


def test_pop_cubature_matches_oracle_and_identities(runtime):
    """FP64 batched cubature of the four kinds follows the oracle's CPU
    restatement step for step at the same budget; J + I1 = I2 within their
    errors; m-Cubes agrees within its error."""
    des = pop_data()
    p, spec, op, ot = _setup(runtime, des)
    igs = [y3.make_integral_volume(p, spec, k, *BOX, (0.3, 0.45, 1.0), eps_rel=1e-9)
           for k in KINDS]
    res = runtime.integrate(igs, y3.options(precision=y3.FP64, max_regions=1024))
    for kind, r in zip(KINDS, res):
        o = O.Result()
        O.lib.y3o_cubature_des(C.byref(op), C.byref(ot.t), O.ptr(des),
                               C.byref(pop_integral(kind, (0.3, 0.45), 1)),
                               C.byref(O.cub_opts(eps_rel=1e-9, max_regions=1024,
                                                  max_eval=1 << 40)), C.byref(o))
        # step for step: the same regions, and values equal to the integrand's
        # device / oracle agreement (1e-8, its exclusion-angle conditioning)
        assert (r["nregions"], r["neval"]) == (o.nregions, o.neval), kind
        assert abs(r["value"] / o.value - 1) < 1e-8, (kind, r["value"], o.value)
    v, e = res["value"], res["error"]
    assert abs(v[1] + v[3] - v[2]) <= e[1] + e[2] + e[3]
    assert np.all(v > 0)
    vg = [y3.Integral.from_buffer_copy(g) for g in igs[:1]]
    vg[0].algorithm, vg[0].max_eval = y3.MCUBES, 4_000_000
    rv = runtime.integrate(vg, y3.options(precision=y3.FP64, vegas_calls=400_000))[0]
    assert abs(rv["value"] - v[0]) <= 3 * rv["error"] + 2 * e[0], (rv, v[0], e[0])


def test_pop_modules_end_to_end(runtime):
    from y3_cluster_cpp_amd import cosmosis as cs
    des = pop_data()
    nc = int(des[11])
    nbm, nbz, nxr, nxz = (int(v) for v in des[12:16])
    o = 16
    cz, cd = des[o:o + nc], des[o + nc:o + 2 * nc]
    o += 2 * nc
    bm, bz = des[o:o + nbm], des[o + nbm:o + nbm + nbz]
    bt = des[o + nbm + nbz:o + nbm + nbz + nbm * nbz].reshape(nbz, nbm)
    o += nbm + nbz + nbm * nbz
    xr, xz = des[o:o + nxr], des[o + nxr:o + nxr + nxz]
    xt = des[o + nxr + nxz:].reshape(nxz, nxr)
    grid = [(0.2, 0.35, 0.0), (0.35, 0.5, 2.0), (0.5, 0.65, 3.0)]
    vals = {}
    for label in ("bSelMargP1GPU", "bSelMargI1GPU", "bSelMargI2GPU", "bSelMargJGPU"):
        cfg = cs.DataBlock()
        for k, v in (("algorithm", "pagani"), ("eps_rel", 1e-3), ("eps_abs", 0.0),
                     ("max_eval", 100_000_000), ("use_cartesian_product", True),
                     ("max_regions", 4096), ("lt_low", BOX[0][0]), ("lt_high", BOX[1][0]),
                     ("zt_low", BOX[0][1]), ("zt_high", BOX[1][1]), ("lnm_low", BOX[0][2]),
                     ("lnm_high", BOX[1][2])):
            cfg.put(label, k, v)
        for i, n in enumerate(("zo_low", "zo_high", "lambda_bin")):
            cfg.put(label, n, np.array([g[i] for g in grid]))
        smp = cs.DataBlock()
        smp.put("cosmological_parameters", "omega_m", 0.3)
        smp.put("cosmological_parameters", "h0", des[0])
        for k, v in zip(("log10_Mmin", "log10_M1", "alpha", "sigma_lambda", "epsilon",
                         "z_pivot"), MOR):
            smp.put("cluster_mor", k, float(v))
        smp.put("distances", "z", np.ascontiguousarray(cz))
        smp.put("distances", "d_c", np.ascontiguousarray(cd))
        smp.put("haloModel", "lnM", np.ascontiguousarray(bm))
        smp.put("haloModel", "z", np.ascontiguousarray(bz))
        smp.put("haloModel", "bias", np.ascontiguousarray(bt))
        smp.put("xi_nl", "r", np.ascontiguousarray(xr))
        smp.put("xi_nl", "z", np.ascontiguousarray(xz))
        smp.put("xi_nl", "xi_nl", np.ascontiguousarray(xt))
        m = cs.Module(label)
        assert m.setup(cfg)
        assert m.execute(smp) == 0
        assert m.cleanup() == 0
        # one volume x three grid points: the cartesian layout's (1, 3)
        v, e = smp.get_array(label, "vals"), smp.get_array(label, "errors")
        assert v.shape == (1, 3)
        vals[label] = (v.ravel(), e.ravel())
    # the library on the same descriptors
    p, spec, _, _ = _setup(runtime, des)
    igs = [y3.make_integral_volume(p, spec, y3.POP_I2, *BOX, g, eps_rel=1e-3) for g in grid]
    r = runtime.integrate(igs, y3.options(precision=y3.FP64, max_regions=4096))
    np.testing.assert_array_equal(r["value"], vals["bSelMargI2GPU"][0])
    (i1, e1), (i2, e2), (j, ej) = (vals[k] for k in ("bSelMargI1GPU", "bSelMargI2GPU",
                                                     "bSelMargJGPU"))
    assert np.all(np.abs(i1 + j - i2) <= e1 + e2 + ej)
