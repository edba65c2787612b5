"""Pins of the oracle's tabulated LC_LT_t (src/models/lc_lt_t.hh:31-63 over the
tables of lc_lt_t.cc:10-214): the reference's SDSS fixture, its unit-mass
test, and table entries read by hand from the source."""
import math

import numpy as np
import pytest
from scipy import integrate

import oracle as O
from conftest import load_columns

L = O.lib


def test_sdss_fixture():
    # test/lc_lt_t.test.cc:11-51: columns z lt lc prob, epsilon 1e-6, margin 1e-12
    d = load_columns("test_lc_lt_t_SDSS.txt")
    assert d.shape == (1000, 4)
    for z, lt, lc, y in d:
        got = L.y3o_lc_lt(lc, lt, z)
        assert abs(got - y) <= max(1e-12, 1e-6 * abs(y)), (z, lt, lc, got, y)


def test_unit_mass():
    # test/lc_lt_probability_unity.test.cc:11-17: int over lc in [0.3, 200]
    # = 1 for lt, zt on 2 x 2 values of [20, 80] x [0.1, 0.3] (tolerance 1e-2,
    # probability_tests.hh)
    for lt in (20.0, 80.0):
        for zt in (0.1, 0.3):
            v, _ = integrate.quad(lambda lc: L.y3o_lc_lt(lc, lt, zt), 0.3, 200.0,
                                  epsabs=0, epsrel=1e-10, limit=400, points=[lt])
            assert abs(v - 1.0) < 1e-2, (lt, zt, v)


def test_table_entries_at_knots():
    """At a knot the clamped bilinear lookup is the table entry itself:
    sigma_arr[0], [1], [22] and fprj / fmsk / mu / tau entries of
    lc_lt_t.cc, read by hand (lt knots 1, 3, ..., zt knots 0.1, 0.15, ...)."""
    p = np.zeros(5)
    L.y3o_lc_lt_params(1.0, 0.1, O.ptr(p))
    assert p[2] == pytest.approx(8.12748540e-01, rel=1e-15)
    L.y3o_lc_lt_params(3.0, 0.1, O.ptr(p))
    assert p[2] == pytest.approx(1.32292795e+00, rel=1e-15)
    L.y3o_lc_lt_params(1.0, 0.15, O.ptr(p))
    assert p[2] == pytest.approx(8.97538569e-01, rel=1e-15)
    # clamps: below the first knots and above the last (lt > 160, zt > 0.3)
    a, b = np.zeros(5), np.zeros(5)
    L.y3o_lc_lt_params(0.2, 0.01, O.ptr(a))
    L.y3o_lc_lt_params(1.0, 0.1, O.ptr(b))
    assert np.array_equal(a, b)
    L.y3o_lc_lt_params(250.0, 1.5, O.ptr(a))
    L.y3o_lc_lt_params(160.0, 0.3, O.ptr(b))
    assert np.array_equal(a, b)
    # bilinear between knots: the cell centre is the mean of its corners
    c = [np.zeros(5) for _ in range(4)]
    for k, (lt, zt) in enumerate([(20., 0.15), (24., 0.15), (20., 0.2), (24., 0.2)]):
        L.y3o_lc_lt_params(lt, zt, O.ptr(c[k]))
    L.y3o_lc_lt_params(22.0, 0.175, O.ptr(a))
    np.testing.assert_allclose(a, sum(c) / 4, rtol=1e-14)


def test_eq15_at_fmsk_zero_is_gauss_plus_emg():
    """With the looked-up parameters, eq. 15 at fmsk = 0 is the two-term
    mixture (1 - fprj) Gauss + fprj EMG (cf. emg_des_t.cuh:172-193)."""
    p = np.zeros(5)
    lt, zt = 33.0, 0.22
    L.y3o_lc_lt_params(lt, zt, O.ptr(p))
    tau, mu, sig, fmsk, fprj = p
    for lc in (20.0, 33.0, 50.0):
        a = L.y3o_lc_lt_eq15(lc, lt, mu, sig, tau, 0.0, fprj)
        b = (1 - fprj) * L.y3o_gaussian(lc, mu, sig) + fprj * L.y3o_emg_pdf(lc, mu, sig, tau)
        assert a == pytest.approx(b, rel=1e-12)
