"""Fixed Gauss-Legendre number counts with a selection function on the
device (y3g_selgl_counts), against the oracle's restatement of the
reference's NumCountsSelGL (n_operator_sel_gl_t.hh), FP64 both sides."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import y3_cluster_cpp_amd as y3

pytestmark = pytest.mark.gpu


def selection_stack(rng, n_bins=12, nm=24, nz=16):
    """A smooth synthetic S_stack [n_bins, n_z, n_lnm]: richness bins as
    sigmoids in lnM, redshift bins as smooth windows in z."""
    lnm = np.linspace(30.0, 35.5, nm)
    z = np.linspace(0.1, 0.8, nz)
    M, Z = np.meshgrid(lnm, z)
    out = np.empty((n_bins, nz, nm))
    for b in range(n_bins):
        m0 = 31.5 + 0.5 * (b % 4) + rng.uniform(-0.1, 0.1)
        zc = 0.25 + 0.15 * (b // 4)
        out[b] = (1 / (1 + np.exp(-(M - m0) / 0.3))) * np.exp(-0.5 * ((Z - zc) / 0.08) ** 2)
    return lnm, z, out


@pytest.fixture(scope="module")
def rt():
    r = y3.Runtime(0)
    yield r
    r.close()


def test_selgl_counts_match_oracle(rt):
    rng = np.random.default_rng(21)
    n_points = 4
    pts = [y3.default_params(omega_m=om, hmf_amp=a) for om, a in
           ((0.3, 0.8), (0.25, 0.75), (0.35, 0.85), (0.28, 0.8))]
    spec = y3.spec_wide()
    rt.tables_build(pts, spec)
    stacks = []
    for k in range(n_points):
        lnm, z, st = selection_stack(rng, n_bins=12 - k)  # ragged bin counts
        rt.selfn_upload(k, lnm, z, st)
        stacks.append((lnm, z, st))
    gl = y3.selgl_spec(30.5, 35.0, 0.15, 0.75, 96, 64)
    got = rt.selgl_counts(gl, n_points, 12)
    for k in range(n_points):
        op = O.default_params(omega_m=pts[k].omega_m, hmf_amp=pts[k].hmf_amp)
        ot = O.OracleTables(op, O.spec_wide())
        lnm, z, st = stacks[k]
        for b in range(12):
            if b >= st.shape[0]:
                assert got[k, b] == 0.0
                continue
            plane = np.ascontiguousarray(st[b].reshape(-1))
            ref = O.lib.y3o_selgl_counts(C.byref(op), C.byref(ot.t), O.ptr(lnm), len(lnm),
                                         O.ptr(z), len(z), O.ptr(plane), 96, 64, 30.5, 35.0,
                                         0.15, 0.75)
            assert abs(got[k, b] / ref - 1) < 1e-12, (k, b, got[k, b], ref)


def test_selgl_node_counts_and_errors(rt):
    rng = np.random.default_rng(5)
    rt.tables_build([y3.default_params()], y3.spec_wide())
    lnm, z, st = selection_stack(rng, n_bins=3)
    rt.selfn_upload(0, lnm, z, st)
    op = O.default_params()
    ot = O.OracleTables(op, O.spec_wide())
    for n_lnm, n_z in ((1, 1), (7, 5), (200, 300)):
        got = rt.selgl_counts(y3.selgl_spec(31.0, 34.0, 0.2, 0.6, n_lnm, n_z), 1, 3)
        plane = np.ascontiguousarray(st[2].reshape(-1))
        ref = O.lib.y3o_selgl_counts(C.byref(op), C.byref(ot.t), O.ptr(lnm), len(lnm), O.ptr(z),
                                     len(z), O.ptr(plane), n_lnm, n_z, 31.0, 34.0, 0.2, 0.6)
        assert abs(got[0, 2] / ref - 1) < 1e-12
    with pytest.raises(RuntimeError):
        rt.selgl_counts(y3.selgl_spec(31.0, 34.0, 0.2, 0.6, 2000, 10), 1, 3)
    with pytest.raises(RuntimeError):
        rt.selgl_counts(y3.selgl_spec(31.0, 34.0, 0.2, 0.6), 2, 3)  # point 1 has no tables
