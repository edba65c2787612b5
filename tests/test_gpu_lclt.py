"""The device LC_LT_t (Costanzi et al. 2018 eq. 15 over the reference's
fitted tables) through the C ABI: the reference's SDSS fixture and the
oracle (tests/test_oracle_lclt.py pins the oracle)."""
import numpy as np
import pytest

import oracle as O
import y3_cluster_cpp_amd as y3
from conftest import load_columns

pytestmark = pytest.mark.gpu


def test_lc_lt_sdss_fixture(runtime):
    # test/lc_lt_t.test.cc:11-51, epsilon 1e-6, margin 1e-12
    d = load_columns("test_lc_lt_t_SDSS.txt")
    got = runtime.eval_functor(y3.F_LC_LT, d[:, [2, 1, 0]])
    assert np.all(np.abs(got - d[:, 3]) <= np.maximum(1e-12, 1e-6 * np.abs(d[:, 3])))


def test_lc_lt_matches_oracle(runtime):
    rng = np.random.default_rng(5)
    n = 2000
    x = np.stack([rng.uniform(0.5, 250.0, n), rng.uniform(0.2, 260.0, n),
                  rng.uniform(0.0, 1.2, n)], 1)
    x[:4] = [[30.0, 0.5, 0.05], [30.0, 300.0, 2.5], [20.0, 20.0, 0.2], [1e-3, 1.0, 0.1]]
    got = runtime.eval_functor(y3.F_LC_LT, x)
    ref = np.array([O.lib.y3o_lc_lt(*row) for row in x])
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-300)
