"""y3g_integrate_noalloc through the C ABI: the same results as y3g_integrate
on two calls in a row (the second smaller), with the call's device memory in
the library's fixed scratch array; inputs beyond its limits are rejected."""
import numpy as np
import pytest

import y3_cluster_cpp_amd as y3
from y3_cluster_cpp_amd.workloads import eps_stream, point_integrals, point_params, prior_points

pytestmark = pytest.mark.gpu


def _batch(rt, n_points, regions):
    spec = y3.spec_wide()
    pts = prior_points(n_points)
    eps = np.minimum(eps_stream(n_points), 1e-4)
    params = [point_params(y3, p) for p in pts]
    rt.tables_build(params, spec)
    igs = []
    for k in range(n_points):
        row = point_integrals(y3, params[k], spec, k, eps[k], y3.PAGANI)
        igs += row[:12] + row[12::9]  # the 12 counts and 20 Delta-Sigma per point
    # four VEGAS integrals of point 0 (5-D counts) at the end
    for i in range(4):
        g = point_integrals(y3, params[0], spec, 0, eps[0], y3.MCUBES)[i]
        igs.append(g)
    return igs, y3.options(precision=y3.FP32SPLIT, max_regions=regions, vegas_calls=20000,
                           vegas_itmax=3)


def _same(a, b, n_cub):
    for f in ("value", "error", "neval", "nregions", "status"):
        assert np.array_equal(a[f][:n_cub], b[f][:n_cub]), f
    # VEGAS histograms accumulate with LDS atomics: equal to rounding
    np.testing.assert_allclose(a["value"][n_cub:], b["value"][n_cub:], rtol=1e-8)
    np.testing.assert_allclose(a["error"][n_cub:], b["error"][n_cub:], rtol=1e-6)


def test_noalloc_matches_integrate_twice():
    rt = y3.Runtime(0)
    try:
        print("noalloc: 3 points, max_regions 1024", flush=True)
        igs, opts = _batch(rt, 3, 1024)
        ref = rt.integrate(igs, opts)
        got = rt.integrate_noalloc(igs, opts)
        _same(ref, got, len(igs) - 4)
        assert np.all(np.isfinite(got["value"])) and np.all(got["neval"] > 0)
        # second call, smaller: the scratch keeps the first call's contents
        small = igs[:7] + igs[-2:]
        print("noalloc: second call, 9 integrals", flush=True)
        ref2 = rt.integrate(small, opts)
        got2 = rt.integrate_noalloc(small, opts)
        _same(ref2, got2, 7)
        ms, nev = rt.last_timing()
        assert nev == int(got2["neval"].sum()) and ms > 0
    finally:
        rt.close()


def test_noalloc_rejects_beyond_limits():
    rt = y3.Runtime(0)
    try:
        igs, opts = _batch(rt, 1, 1024)
        big = y3.options(precision=y3.FP32SPLIT, max_regions=1 << 20)
        with pytest.raises(RuntimeError, match="scratch"):
            rt.integrate_noalloc(igs[:2], big)
        with pytest.raises(RuntimeError, match="single"):
            rt.integrate_noalloc(igs[:2], y3.options(precision=y3.FP64, single=1,
                                                     max_regions=1024))
        # within limits again afterwards
        ref = rt.integrate(igs[:3], opts)
        got = rt.integrate_noalloc(igs[:3], opts)
        _same(ref, got, 3)
    finally:
        rt.close()
