"""Pins of the oracle's tabulated miscentred NFW profile (NFW_DSIGMA_MIS,
src/models/nfw_dsigma_mis.cuh): the table's axes against the reference's
files, its entries against an independent adaptive quadrature of the same
profile, and the lookup against the cluster_toolkit values of
test/nfw_dsigma_mis.test.cu."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from conftest import load_columns

L = O.lib
RHOC = 2.77533742639e11


def axes():
    a = np.zeros(1000)
    b = np.zeros(250)
    L.y3o_nfw_mis_axes(O.ptr(a), O.ptr(b))
    return a, b


def halo(c, rho, lnM):
    r_s = (3 * math.exp(lnM) / (800 * math.pi * rho)) ** (1 / 3) / c
    delta_c = (200 * c ** 3 / 3) / (math.log(1 + c) - c / (1 + c))
    return r_s, delta_c


def test_axes_are_the_reference_files():
    a, b = axes()
    ra = load_columns("table_1000_1e-03_5e+03_gamma_logx.txt")[:, 0]
    rb = load_columns("table_1000_1e-03_5e+03_gamma_logxmis.txt")[:, 0]
    assert a.shape == ra.shape and b.shape == rb.shape
    assert np.abs(a - ra).max() < 1e-14 and np.abs(b - rb).max() < 1e-14


def test_entries_match_adaptive_quadrature():
    """G(x, xmis) by the fixed graded rule against y3o_dsigma_mis_gamma
    (nested adaptive GK21 over the gamma kernel and the ring), converted to
    units of rho_s r_s; where the ring average cancels to below 1e-10 of
    g both carry rounding, so there the bound is 1e-4."""
    c, rho, lnM = 4.0, 0.3 * RHOC, 32.2361913
    r_s, dc = halo(c, rho, lnM)
    for x in (1e-3, 0.013, 0.1, 0.37, 1.0, 2.9, 10.0, 77.0, 600.0, 5e3):
        for m in (1e-2, 0.05, 0.3, 1.0, 4.0, 30.0, 5e2):
            fixed = L.y3o_nfw_mis_unit(x, m)
            adapt = L.y3o_dsigma_mis_gamma(x * r_s, m * r_s, lnM, rho, c, 1e-9) / (
                1e-12 * dc * rho * r_s)
            tol = 1e-7 if adapt > 1e-10 else 1e-4
            assert fixed == pytest.approx(adapt, rel=tol), (x, m)
            assert fixed > 0.0
    # the entry is ln(G / 2)
    assert L.y3o_nfw_mis_entry(math.log(0.5), math.log(0.2)) == pytest.approx(
        math.log(0.5 * L.y3o_nfw_mis_unit(0.5, 0.2)), rel=1e-15)


def test_small_offset_limit_is_centred_profile():
    """xmis -> 0: the gamma-kernel average tends to the centred g(x)."""
    for x in (0.05, 0.5, 3.0):
        assert L.y3o_nfw_mis_unit(x, 1e-5) == pytest.approx(L.y3o_nfw_g(x), rel=2e-4)


def _local_lookup(r, rmis, lnM, c, rho):
    """NFW_DSIGMA_MIS through the four table knots around (ln x, ln xmis)."""
    a, b = axes()
    r_s, _ = halo(c, rho, lnM)
    lx, lm = math.log(r / r_s), math.log(rmis / r_s)
    i = int(np.clip(np.searchsorted(a, lx) - 1, 0, 998))
    j = int(np.clip(np.searchsorted(b, lm) - 1, 0, 248))
    ax, bx = np.ascontiguousarray(a[i:i + 2]), np.ascontiguousarray(b[j:j + 2])
    T = np.array([L.y3o_nfw_mis_entry(ax[p], bx[q]) for q in (0, 1) for p in (0, 1)])
    return L.y3o_nfw_dsigma_mis_lookup(O.ptr(ax), 2, O.ptr(bx), 2, O.ptr(T), r, rmis, lnM, c,
                                       rho)


# test/nfw_dsigma_mis.test.cu:20-66: c = 4, rho = 0.3 rho_crit, lnM =
# 32.2361913, 10 radii x 3 offsets, cluster_toolkit answers, epsilon 5e-3
CT_R = [0.1, 0.16681005, 0.27825594, 0.46415888, 0.77426368, 1.29154967,
        2.15443469, 3.59381366, 5.9948425, 10.]
CT_RMIS = [0.1, 0.5, 2.0]
CT_VALUES = [12.3685326, 0.3899747, 0.0104446, 22.7491859, 0.9300333, 0.0264995,
             33.3082693, 2.0334094, 0.0641354, 36.0917619, 3.9092169, 0.1455426,
             28.3882168, 6.2393766, 0.3031914, 17.5178459, 7.6804544, 0.5638920,
             9.4010319, 6.7788772, 0.9002713, 4.6358405, 4.2448975, 1.1650680,
             2.1556775, 2.1118245, 1.1349319, 0.9599990, 0.9545601, 0.7860161]


def test_lookup_reproduces_cluster_toolkit():
    for i, r in enumerate(CT_R):
        for j, rm in enumerate(CT_RMIS):
            got = _local_lookup(r, rm, 32.2361913, 4.0, 0.3 * RHOC)
            exp = CT_VALUES[j + 3 * i]
            assert abs(got - exp) <= max(1e-12, 5e-3 * abs(exp)), (r, rm, got, exp)


def test_lookup_normalisation_at_knots():
    """At a knot the lookup is the entry itself: 2 r_s delta_c rho e^T 1e-12
    equals the adaptive physical profile (pins the normalisation)."""
    a, b = axes()
    c, rho, lnM = 5.0, 0.25 * RHOC, 33.0
    r_s, _ = halo(c, rho, lnM)
    for i, j in [(400, 100), (520, 130), (610, 60)]:
        r, rm = math.exp(a[i]) * r_s, math.exp(b[j]) * r_s
        T = np.array([L.y3o_nfw_mis_entry(a[i], b[j])] * 4)
        ax, bx = np.ascontiguousarray(a[i:i + 2]), np.ascontiguousarray(b[j:j + 2])
        got = L.y3o_nfw_dsigma_mis_lookup(O.ptr(ax), 2, O.ptr(bx), 2, O.ptr(T), r, rm, lnM, c,
                                          rho)
        ref = L.y3o_dsigma_mis_gamma(r, rm, lnM, rho, c, 1e-9)
        assert got == pytest.approx(ref, rel=1e-7)


def test_lookup_is_clamped_bilinear():
    """Bilinear in (ln x, ln xmis) inside, clamped to the edges outside."""
    ax = np.ascontiguousarray([0.0, 1.0, 3.0])
    bx = np.ascontiguousarray([-1.0, 2.0])
    T = np.ascontiguousarray([[1.0, 2.0, 4.0], [0.5, -1.0, 7.0]]).reshape(-1)
    c, rho, lnM = 4.0, 0.3 * RHOC, 32.0
    r_s, dc = halo(c, rho, lnM)
    norm = 2 * r_s * dc * rho * 1e-12

    def look(lx, lm):
        return L.y3o_nfw_dsigma_mis_lookup(O.ptr(ax), 3, O.ptr(bx), 2, O.ptr(T),
                                           math.exp(lx) * r_s, math.exp(lm) * r_s, lnM, c, rho)
    # (0.5, 0.5): t = 0.5, u = 0.5 -> mean of 1, 2, 0.5, -1
    assert look(0.5, 0.5) == pytest.approx(norm * math.exp(0.625), rel=1e-12)
    assert look(-5.0, -9.0) == pytest.approx(norm * math.exp(1.0), rel=1e-12)
    assert look(9.0, 9.0) == pytest.approx(norm * math.exp(7.0), rel=1e-12)
