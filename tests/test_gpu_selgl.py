"""Fixed Gauss-Legendre number counts with a selection function on the
device (y3g_selgl_counts), against the oracle's restatement of the
reference's NumCountsSelGL (n_operator_sel_gl_t.hh), FP64 both sides."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import y3_cluster_cpp_amd as y3
from y3_cluster_cpp_amd import workloads

pytestmark = pytest.mark.gpu


def selection_stack(rng, n_bins=12, nm=24, nz=16):
    return workloads.selection_stack(rng, n_bins, nm, nz)


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


def test_numcountssel_module_matches_oracle():
    """The NumCountsSel CosmoSIS module (setup / execute / cleanup through
    the C entry points): numcountssel/vals per bin_index equal the oracle on
    the sample's selection function, after a parameter change too, and a
    bin_index outside S_stack is an execute error (the reference's
    NumCountsSelGL::evaluate throws out_of_range)."""
    from y3_cluster_cpp_amd import cosmosis as cs
    L = "NumCountsSel"
    rng = np.random.default_rng(8)
    lnm, z, st = selection_stack(rng, n_bins=6)
    bins = [0, 3, 5, 1, 3]
    cfg = cs.DataBlock()
    for k, v in (("n_lnm", 64), ("n_z", 48), ("zt_low", 0.15), ("zt_high", 0.75),
                 ("lnm_low", 30.5), ("lnm_high", 35.0)):
        cfg.put(L, k, v)
    cfg.put(L, "bin_index", np.asarray(bins, dtype=float))
    m = cs.Module(L)
    assert m.setup(cfg), "setup returned NULL"
    for om in (0.3, 0.27):
        smp = cs.DataBlock()
        smp.put("cosmological_parameters", "omega_m", om)
        smp.put("sel_function", "lnM", lnm)
        smp.put("sel_function", "z", z)
        smp.put("sel_function", "S_stack", st)
        assert m.execute(smp) == 0
        vals = smp.get_array("numcountssel", "vals")
        assert vals.shape == (len(bins),)
        op = O.default_params(omega_m=om)
        ot = O.OracleTables(op, O.spec_wide())
        for i, b in enumerate(bins):
            plane = np.ascontiguousarray(st[b].reshape(-1))
            ref = O.lib.y3o_selgl_counts(C.byref(op), C.byref(ot.t), O.ptr(lnm), len(lnm),
                                         O.ptr(z), len(z), O.ptr(plane), 64, 48, 30.5, 35.0,
                                         0.15, 0.75)
            assert abs(vals[i] / ref - 1) < 1e-12, (om, b, vals[i], ref)
    smp = cs.DataBlock()
    smp.put("cosmological_parameters", "omega_m", 0.3)
    smp.put("sel_function", "lnM", lnm)
    smp.put("sel_function", "z", z)
    smp.put("sel_function", "S_stack", st[:5])  # bin 5 no longer exists
    assert m.execute(smp) != 0
    assert m.cleanup() == 0


def test_selgl_nodes_outside_axes_nonuniform_axes(rt):
    """GL windows wider than the selection axes (nodes outside them add
    nothing) and non-uniform axes (the bracket search), against the oracle."""
    rng = np.random.default_rng(13)
    rt.tables_build([y3.default_params(omega_m=0.32)], y3.spec_wide())
    lnm = np.sort(np.concatenate([[30.2, 35.1], rng.uniform(30.2, 35.1, 29)]))
    z = np.sort(np.concatenate([[0.12, 0.7], rng.uniform(0.12, 0.7, 11)]))
    st = rng.uniform(0.0, 1.0, (5, len(z), len(lnm)))
    rt.selfn_upload(0, lnm, z, st)
    op = O.default_params(omega_m=0.32)
    ot = O.OracleTables(op, O.spec_wide())
    win = (29.0, 36.0, 0.05, 0.9)
    got = rt.selgl_counts(y3.selgl_spec(*win, 96, 64), 1, 5)
    for b in range(5):
        plane = np.ascontiguousarray(st[b].reshape(-1))
        ref = O.lib.y3o_selgl_counts(C.byref(op), C.byref(ot.t), O.ptr(lnm), len(lnm),
                                     O.ptr(z), len(z), O.ptr(plane), 96, 64, *win)
        assert abs(got[0, b] / ref - 1) < 1e-12, (b, got[0, b], ref)


def test_selgl_long_axes_searched_in_global(rt):
    """Selection axes longer than the LDS staging size (here 900 + 200
    doubles) are searched in global memory; two points with different grids
    in one launch, against the oracle."""
    rng = np.random.default_rng(17)
    rt.tables_build([y3.default_params(), y3.default_params(omega_m=0.26)], y3.spec_wide())
    stacks = [workloads.selection_stack(rng, 4, 900, 200),
              workloads.selection_stack(rng, 3, 30, 12)]
    for k, (lnm, z, st) in enumerate(stacks):
        rt.selfn_upload(k, lnm, z, st)
    got = rt.selgl_counts(y3.selgl_spec(30.5, 35.0, 0.15, 0.75, 96, 64), 2, 4)
    for k, op in enumerate((O.default_params(), O.default_params(omega_m=0.26))):
        ot = O.OracleTables(op, O.spec_wide())
        lnm, z, st = stacks[k]
        for b in range(st.shape[0]):
            plane = np.ascontiguousarray(st[b].reshape(-1))
            ref = O.lib.y3o_selgl_counts(C.byref(op), C.byref(ot.t), O.ptr(lnm), len(lnm),
                                         O.ptr(z), len(z), O.ptr(plane), 96, 64, 30.5, 35.0,
                                         0.15, 0.75)
            assert abs(got[k, b] / ref - 1) < 1e-12, (k, b, got[k, b], ref)
        assert np.all(got[k, st.shape[0]:] == 0.0)
