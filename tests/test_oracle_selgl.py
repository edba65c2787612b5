"""Pins of the oracle's fixed Gauss-Legendre number counts with a selection
function (y3o_selgl.c; reference n_operator_sel_gl_t.hh NumCountsSelGL,
sel_function_t.hh, omega_z_des.hh)."""
import ctypes as C
import warnings

import numpy as np
from scipy import integrate

import oracle as O

L = O.lib


def test_omega_z_des_pieces():
    # hand evaluation of each piece from the reference's coefficients
    # (leading power first): z < 0.504, 0.504 <= z < 0.7 (at z - 0.6), z >= 0.7
    z = 0.3
    assert abs(L.y3o_omega_z_des(z) - (-0.00262353 * z ** 2 + 0.01940118 * z + 0.45133063)) < 1e-15
    x = 0.65 - 0.6
    ref = (1.33647377e4 * x ** 5 + 1.35291046e3 * x ** 4 - 1.26204891e2 * x ** 3
           - 2.83454918e1 * x ** 2 - 2.26465905 * x + 3.84958753e-1)
    assert abs(L.y3o_omega_z_des(0.65) - ref) < 1e-14
    z = 0.8
    ref = -1.88101967 * z ** 3 + 4.8071839 * z ** 2 - 4.11424324 * z + 1.18196785
    assert abs(L.y3o_omega_z_des(z) - ref) < 1e-14
    assert L.y3o_omega_z_des(0.6) == 3.84958753e-1


def _bilinear_table(nm=9, nz=7):
    lnm = np.linspace(30.0, 35.0, nm)
    z = np.linspace(0.1, 0.8, nz)
    M, Z = np.meshgrid(lnm, z)  # (nz, nm), lnM fastest
    plane = 0.3 + 0.02 * (M - 30) + 0.5 * Z + 0.04 * (M - 30) * Z
    return lnm, z, np.ascontiguousarray(plane.reshape(-1))


def test_sel_function_bilinear_and_zero_outside():
    lnm, z, plane = _bilinear_table()
    f = lambda a, b: L.y3o_sel_function(O.ptr(lnm), len(lnm), O.ptr(z), len(z), O.ptr(plane), a, b)
    rng = np.random.default_rng(3)
    for _ in range(200):
        a, b = rng.uniform(30, 35), rng.uniform(0.1, 0.8)
        exact = 0.3 + 0.02 * (a - 30) + 0.5 * b + 0.04 * (a - 30) * b
        assert abs(f(a, b) - exact) < 1e-13
    assert f(29.99, 0.5) == 0.0 and f(35.01, 0.5) == 0.0
    assert f(32.0, 0.09) == 0.0 and f(32.0, 0.81) == 0.0
    assert abs(f(30.0, 0.1) - plane[0]) < 1e-15


def test_gl_nodes_ab_exact():
    for n in (3, 8, 20):
        x = np.zeros(n)
        w = np.zeros(n)
        L.y3o_gl_nodes_ab(n, 0.2, 1.7, O.ptr(x), O.ptr(w))
        for deg in range(2 * n):
            exact = (1.7 ** (deg + 1) - 0.2 ** (deg + 1)) / (deg + 1)
            assert abs(np.dot(w, x ** deg) - exact) <= 1e-12 * max(1.0, abs(exact))


def _setup():
    p = O.default_params()
    t = O.OracleTables(p, O.spec_wide())
    lnm, z, plane = _bilinear_table(16, 12)
    return p, t, lnm, z, plane


def _counts(p, t, lnm, z, plane, n_lnm, n_z, box):
    return L.y3o_selgl_counts(C.byref(p), C.byref(t.t), O.ptr(lnm), len(lnm), O.ptr(z), len(z),
                              O.ptr(plane), n_lnm, n_z, *box)


BOX = (30.5, 34.5, 0.2, 0.65)


def test_selgl_counts_linear_in_selection():
    p, t, lnm, z, plane = _setup()
    a = _counts(p, t, lnm, z, plane, 24, 16, BOX)
    b = _counts(p, t, lnm, z, np.ascontiguousarray(2.5 * plane), 24, 16, BOX)
    assert abs(b / a - 2.5) < 1e-14


def test_selgl_counts_against_adaptive_integral():
    """The GL double sum against scipy's adaptive 2-D integral of the same
    integrand assembled from the separately pinned pieces (HMF and dV/dz of
    the point's tables, Omega(z), the selection function): a dropped factor,
    a wrong weight or node map fails."""
    p, t, lnm, z, plane = _setup()

    def integrand(zt, lnM):
        return (L.y3o_dvdz(C.byref(p), C.byref(t.t), zt) * L.y3o_omega_z_des(zt) *
                L.y3o_hmf(C.byref(p), C.byref(t.t), lnM, zt) *
                L.y3o_sel_function(O.ptr(lnm), len(lnm), O.ptr(z), len(z), O.ptr(plane), lnM, zt))

    # split at the piecewise-polynomial break of Omega(z) (z = 0.504)
    ref = 0.0
    with warnings.catch_warnings():  # the tables' C0 knots slow the adaptive rule
        warnings.simplefilter("ignore", integrate.IntegrationWarning)
        for zlo, zhi in ((BOX[2], 0.504), (0.504, BOX[3])):
            v, _ = integrate.dblquad(integrand, BOX[0], BOX[1], zlo, zhi, epsabs=0, epsrel=1e-6)
            ref += v
    gl = _counts(p, t, lnm, z, plane, 96, 64, BOX)
    fine = _counts(p, t, lnm, z, plane, 400, 300, BOX)
    assert abs(gl / ref - 1) < 2e-3      # the reference's default node counts
    assert abs(fine / ref - 1) < 1e-4    # converging to the adaptive value
