"""GPU: the request server (y3g_server_* of include/y3g.h).

- a request's results are bit-identical whether it is served alone or in
  a mix of other requests (different parameter points and tolerances), and
  equal to the library's batched call on the same descriptors;
- the LRU cache answers a repeated request without the device;
- admission: a budget below one lane's region slab is refused, a budget
  for one lane admits one;
- a burst of mock-MCMC requests completes, every result checked against
  the FP64 oracle's golden values within its eps_rel (capped at 1e-4 for
  the count) with a reported error that bounds the true error.
"""
import json
import os

import numpy as np
import pytest

import y3_cluster_cpp_amd as y3
from conftest import GOLDEN
from y3_cluster_cpp_amd import workloads as W

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "golden_points.json")) as f:
    GP = json.load(f)
PTS = np.array(GP["points"])
REF = np.concatenate([np.array(GP["counts"]), np.array(GP["dsigma"])], axis=1)
EPS = W.eps_stream(len(PTS))


def request(k, counts_only, eps_scale=1.0, cap=1e-3):
    p = W.point_params(y3, PTS[k])
    eps = np.minimum(EPS[k] * eps_scale, cap)
    return p, W.point_integrals(y3, p, y3.spec_wide(), 0, eps, y3.PAGANI, counts_only)


def opts(**kw):
    o = y3.server_options(**kw)
    o.options.max_regions = 8192
    return o


def test_alone_vs_mix_bit_identical(runtime):
    srv = y3.Server(0, opts(lanes=2, cache_entries=0, max_batch_integrals=256))
    p, igs = request(3, counts_only=False)
    sub = igs[:12] + igs[12::20]           # 12 counts + 9 Delta-Sigma
    t = srv.submit(p, sub)
    alone, lat = srv.wait(t)
    assert lat > 0
    tickets = []
    for k, full in [(5, True), (7, False), (3, None), (9, True), (11, False)]:
        if full is None:
            tickets.append(("A", srv.submit(p, sub)))
        else:
            pk, ik = request(k, counts_only=not full, eps_scale=0.5)
            tickets.append((k, srv.submit(pk, ik)))
    outs = {k: srv.wait(t)[0] for k, t in tickets}
    st = srv.stats()
    srv.close()
    mixed = outs["A"]
    for f in ("value", "error", "nregions", "neval", "status"):
        np.testing.assert_array_equal(mixed[f], alone[f], err_msg=f)
    assert st["batches"] >= 2 and st["requests_done"] == 6
    # the same descriptors through the library's batched call
    runtime.tables_build(p, y3.spec_wide())
    direct = runtime.integrate(sub, y3.options(max_regions=8192))
    np.testing.assert_array_equal(direct["value"], alone["value"])


def test_cache_hits_and_admission():
    srv = y3.Server(0, opts(lanes=1, cache_entries=1000))
    p, igs = request(1, counts_only=True)
    a, _ = srv.wait(srv.submit(p, igs))
    b, lat = srv.wait(srv.submit(p, igs))
    st = srv.stats()
    srv.close()
    np.testing.assert_array_equal(a, b)
    assert st["cache_hits"] == 12 and st["integrals_computed"] == 12 and st["batches"] == 1
    assert lat < 0.05
    o = opts(lanes=4, hbm_budget_bytes=1 << 20)
    with pytest.raises(RuntimeError, match="over the budget"):
        y3.Server(0, o)
    probe = y3.Server(0, opts(lanes=1))
    per_lane = probe.stats()["slab_bytes_per_lane"]
    probe.close()
    assert per_lane > 0
    srv = y3.Server(0, opts(lanes=4, hbm_budget_bytes=int(per_lane * 1.5)))
    assert srv.stats()["lanes"] == 1
    srv.close()


def test_mcmc_burst_against_golden():
    srv = y3.Server(0, opts(lanes=2, cache_entries=0))
    t_arr, reqs = W.mcmc_requests(PTS, 24, rate=1000.0)
    sent = []
    for k, full in reqs:
        p, igs = request(k, counts_only=not full, cap=1e-3)
        sent.append((k, full, srv.submit(p, igs)))
    lat = []
    for k, full, t in sent:
        out, l = srv.wait(t)
        lat.append(l)
        n = len(out)
        ref = REF[k, :n]
        eps = np.minimum(EPS[k, :n], 1e-3)
        err = np.abs(out["value"] - ref)
        assert np.all(out["error"] >= err), (k, np.where(out["error"] < err))
        conv = out["status"] == 0    # converged ones meet their tolerance
        assert np.all(err[conv] <= eps[conv] * np.abs(ref[conv]))
        assert np.all(err <= 1e-2 * np.abs(ref))
    srv.close()
    assert len(lat) == 24 and min(lat) > 0
