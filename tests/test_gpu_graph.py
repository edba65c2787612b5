"""GPU: one batched step captured as a HIP graph (y3g_graph_*).

Replayed several times with new inputs, with the program's own eager work
and synchronisations in between, every replay equals the eager
y3g_tables_build + y3g_integrate on the same inputs bit for bit; replaying
the first inputs again repeats the first results; a reallocation of the
runtime's region slab invalidates the graph (an error, not a stale
address)."""
import numpy as np
import pytest

import y3_cluster_cpp_amd as y3
from y3_cluster_cpp_amd import workloads as W

pytestmark = pytest.mark.gpu

NP, PER = 3, 32          # 3 points x (12 counts + 20 Delta-Sigma)


def step_inputs(first):
    pts = W.prior_points(first + NP)[first:]
    eps = np.minimum(W.eps_stream(first + NP)[first:], 1e-3)
    params = [W.point_params(y3, pt) for pt in pts]
    igs = []
    for k, p in enumerate(params):
        full = W.point_integrals(y3, p, y3.spec_wide(), k, eps[k], y3.PAGANI)
        igs += full[:12] + full[12::9][:20]
    assert len(igs) == NP * PER
    return params, igs


def test_graph_replays_equal_eager(runtime):
    spec = y3.spec_wide()
    o = y3.options(precision=y3.FP32SPLIT, max_regions=8192, bands=[(1e-4, 8192), (1e-3, 4096)])
    g = y3.Graph(runtime, spec, NP, NP * PER, o)
    try:
        _replays(runtime, spec, o, g)
    finally:
        g.close()


def _replays(runtime, spec, o, g):
    outs = []
    for first in (0, 7, 0):
        params, igs = step_inputs(first)
        for k, p in enumerate(params):
            g.params[k] = p
        for i, ig in enumerate(igs):
            g.integrals[i] = ig
        g.run()
        got = g.results()
        ms, nev = runtime.last_timing()
        assert ms > 0 and nev == int(got["neval"].sum())
        # the same step eagerly, on the same runtime, between the replays
        runtime.tables_build(params, spec)
        eager = runtime.integrate(igs, o)
        for f in ("value", "error", "nregions", "neval", "status"):
            np.testing.assert_array_equal(got[f], eager[f], err_msg=f)
        outs.append(got)
    np.testing.assert_array_equal(outs[0], outs[2])
    assert not np.array_equal(outs[0]["value"], outs[1]["value"])
    # a larger region budget reallocates the runtime's slab: the graph refuses
    params, igs = step_inputs(0)
    runtime.tables_build(params, spec)
    runtime.integrate(igs[:4], y3.options(max_regions=16384))
    with pytest.raises(RuntimeError, match="reallocated"):
        g.run()


def test_graph_rejects_bad_inputs(runtime):
    spec = y3.spec_wide()
    g = y3.Graph(runtime, spec, 1, 2, y3.options(max_regions=1024))
    params, igs = step_inputs(0)
    g.params[0] = params[0]
    g.integrals[0] = igs[0]
    g.integrals[1] = igs[PER]        # point 1: out of range for a 1-point graph
    try:
        with pytest.raises(RuntimeError, match="malformed"):
            g.run()
    finally:
        g.close()
    with pytest.raises(RuntimeError):
        y3.Graph(runtime, spec, 1, 2, y3.options(single=1))
