"""NFW_DSIGMA_MIS's table built on the device and the device functor over it,
against the oracle (tests/test_oracle_nfwtab.py pins the oracle), through the
C ABI."""
import math

import numpy as np
import pytest

import oracle as O
import y3_cluster_cpp_amd as y3
from conftest import load_columns
from test_oracle_nfwtab import CT_R, CT_RMIS, CT_VALUES, RHOC

pytestmark = pytest.mark.gpu
L = O.lib


@pytest.fixture(scope="module")
def table(runtime):
    print("building the NFW_DSIGMA_MIS table on the device", flush=True)
    runtime.nfw_mis_table_build()
    return runtime.nfw_mis_table_download()


def test_table_axes_and_entries(table):
    a, b, T = table
    assert a.shape == (1000,) and b.shape == (250,) and T.shape == (250, 1000)
    assert np.abs(a - load_columns("table_1000_1e-03_5e+03_gamma_logx.txt")[:, 0]).max() < 1e-14
    assert np.abs(b - load_columns("table_1000_1e-03_5e+03_gamma_logxmis.txt")[:, 0]).max() < 1e-14
    assert np.all(np.isfinite(T))
    rng = np.random.default_rng(7)
    idx = [(0, 0), (0, 999), (249, 0), (249, 999), (125, 500)]
    idx += [(int(j), int(i)) for j, i in zip(rng.integers(0, 250, 300), rng.integers(0, 1000, 300))]
    for j, i in idx:
        ref = L.y3o_nfw_mis_entry(a[i], b[j])
        # same rule, summed in another order: equal to the rounding of its
        # terms, which the ring average cancels by the factor
        # sum |terms| / G (large where xmis >> x); bound in T = ln(G / 2)
        x, m = math.exp(a[i]), math.exp(b[j])
        cond = L.y3o_nfw_mis_unit_abs(x, m) / L.y3o_nfw_mis_unit(x, m)
        tol = 1e-12 + 256 * 2.2e-16 * cond
        assert abs(T[j, i] - ref) < tol, (j, i, T[j, i], ref, cond)


def test_functor_reproduces_cluster_toolkit(runtime, table):
    # test/nfw_dsigma_mis.test.cu:20-66 (c = 4, rho = 0.3 rho_crit, 5e-3)
    x = np.array([[r, rm, 32.2361913] for r in CT_R for rm in CT_RMIS])
    got = runtime.eval_functor(y3.F_NFW_DSIGMA_MIS, x, [4.0, 0.3 * RHOC])
    exp = np.array(CT_VALUES)
    assert np.all(np.abs(got - exp) <= np.maximum(1e-12, 5e-3 * np.abs(exp)))


def test_functor_matches_oracle_lookup(runtime, table):
    a, b, T = table
    aa, bb, tt = (np.ascontiguousarray(v).reshape(-1) for v in (a, b, T))
    rng = np.random.default_rng(11)
    n = 400
    r = 10 ** rng.uniform(-3, 2.5, n)
    rm = 10 ** rng.uniform(-3, 1.5, n)
    lnM = rng.uniform(29.9, 35.7, n)
    r[:3], rm[:3] = [1e-9, 1e6, 0.3], [1e-9, 1e6, 1e7]  # clamps at the edges
    for c, rho in [(4.0, 0.3 * RHOC), (6.5, 0.22 * RHOC)]:
        got = runtime.eval_functor(y3.F_NFW_DSIGMA_MIS, np.stack([r, rm, lnM], 1), [c, rho])
        ref = np.array([L.y3o_nfw_dsigma_mis_lookup(O.ptr(aa), 1000, O.ptr(bb), 250, O.ptr(tt),
                                                    r[k], rm[k], lnM[k], c, rho)
                        for k in range(n)])
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)


def test_upload_download_round_trip(runtime, table):
    a, b, T = table
    rt = y3.Runtime(0)
    try:
        with pytest.raises(RuntimeError, match="table"):
            rt.eval_functor(y3.F_NFW_DSIGMA_MIS, [[1.0, 0.2, 32.0]], [4.0, 0.3 * RHOC])
        with pytest.raises(RuntimeError, match="increasing"):
            rt.nfw_mis_table_upload(a[::-1], b, T)
        rt.nfw_mis_table_upload(a, b, T)
        a2, b2, T2 = rt.nfw_mis_table_download()
        assert np.array_equal(a, a2) and np.array_equal(b, b2) and np.array_equal(T, T2)
        x = [[0.5, 0.1, 32.5], [2.0, 1.0, 34.0]]
        assert np.array_equal(rt.eval_functor(y3.F_NFW_DSIGMA_MIS, x, [4.0, 0.3 * RHOC]),
                              runtime.eval_functor(y3.F_NFW_DSIGMA_MIS, x, [4.0, 0.3 * RHOC]))
    finally:
        rt.close()
