"""The reference's Sigma_crit^-1 lookups on the device, through the C ABI:
the 1-D average_sigma_crit_inv/{zlense, sci_average} table
(lensing_weights.hh:55-60, read with clamp(zt) at :308) and the (r, z)
sigmaCritInv/sigma_crit_inv table of GAMMA_1H_NFW (gamma_1h_nfw.cuh:81-88,
read with clamp(r, zt) at :120), against the oracle's GSL-semantics
interpolators (pinned by the interp_2d.test.cc cases) and closed forms."""
import numpy as np
import pytest

import oracle as O
import y3_cluster_cpp_amd as y3

pytestmark = pytest.mark.gpu


def test_sci_average_table(runtime):
    rng = np.random.default_rng(3)
    z = np.sort(rng.uniform(0.05, 1.2, 40))
    v = rng.uniform(1e-5, 4e-4, 40)
    par = np.concatenate([[40], z, v])
    q = np.concatenate([rng.uniform(0.0, 1.4, 500), z[:5], [-1.0, 9.0]])
    got = runtime.eval_functor(y3.F_SCI_AVERAGE, q[:, None], par)
    ref = np.array([O.lib.y3o_interp1d_clamp(O.ptr(z), O.ptr(v), 40, x) for x in q])
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=0)
    # a linear table is reproduced exactly inside, its end values outside
    zl = np.linspace(0.1, 0.9, 9)
    par = np.concatenate([[9], zl, 2.0 * zl + 1.0])
    got = runtime.eval_functor(y3.F_SCI_AVERAGE, np.array([[0.33], [0.0], [2.0]]), par)
    np.testing.assert_allclose(got, [1.66, 1.2, 2.8], rtol=1e-14)


def test_sci_rz_table(runtime):
    rng = np.random.default_rng(4)
    r = np.sort(rng.uniform(0.05, 30.0, 25))
    z = np.sort(rng.uniform(0.1, 0.9, 12))
    t = rng.uniform(1e-5, 4e-4, (12, 25))  # shape (n_z, n_r): r fastest
    par = np.concatenate([[25, 12], r, z, t.reshape(-1)])
    q = np.stack([rng.uniform(0.0, 35.0, 600), rng.uniform(0.0, 1.0, 600)], 1)
    got = runtime.eval_functor(y3.F_SCI_RZ, q, par)
    tt = np.ascontiguousarray(t.reshape(-1))
    ref = np.array([O.lib.y3o_interp2d_clamp(O.ptr(r), 25, O.ptr(z), 12, O.ptr(tt), a, b)
                    for a, b in q])
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=0)
    # at a knot: the entry itself
    got = runtime.eval_functor(y3.F_SCI_RZ, np.array([[r[7], z[3]]]), par)
    assert got[0] == pytest.approx(t[3, 7], rel=1e-15)
    with pytest.raises(RuntimeError, match="SCI_RZ"):
        runtime.eval_functor(y3.F_SCI_RZ, q, par[:-1])
