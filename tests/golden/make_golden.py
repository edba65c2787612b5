"""Golden values of the synthetic workloads from the FP64 oracle.

    python tests/golden/make_golden.py            # configs 0-2 (fiducial)
    python tests/golden/make_golden.py --points N # + first N parameter
                                                  #   points of the config-3
                                                  #   prior-box stream
    python tests/golden/make_golden.py --stream 4096 --stride 16
                                                  # + every 16th point of the
                                                  #   4096-point stream

Outputs tests/golden/golden_fiducial.json, golden_points.json (the first N
points) and golden_stream.json (the stratified subset: point indices and
their 192 values each).  The
golden integrator (oracle/y3o_quad.c y3o_golden) reduces the lambda_obs and
z_obs integrals to closed forms and integrates the rest on knot-bounded
Gauss-Legendre panels; it is pinned against brute-force cubature in
tests/test_oracle_quadrature.py.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
LO_EDGES = [20.0, 30.0, 45.0, 60.0, 300.0]
ZO_EDGES = [0.20, 0.35, 0.50, 0.65]
RADII = list(np.logspace(np.log10(0.1), np.log10(30.0), 15))


def params_dict(p):
    return {n: getattr(p, n) for n, _ in O.Params._fields_}


def prior_points(n, seed=0x59334331):
    """Parameter points of the config-3 prior box (same generator as bench.py)."""
    rng = np.random.Generator(np.random.Philox(key=seed))
    u = rng.random((n, 5))
    lo = np.array([0.2, 0.7, 0.7, 0.1, 0.0])
    hi = np.array([0.4, 0.9, 1.1, 0.4, 0.4])
    return lo + u * (hi - lo)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--points", type=int, default=0)
    ap.add_argument("--stream", type=int, default=0)
    ap.add_argument("--stride", type=int, default=16)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--skip-fiducial", action="store_true")
    a = ap.parse_args()
    if not a.skip_fiducial:
        fiducial()
    if a.points > 0:
        pts = prior_points(a.points)
        res = batch_golden(pts, a.threads)
        with open(os.path.join(HERE, "golden_points.json"), "w") as f:
            json.dump({"seed": 0x59334331, "points": pts.tolist(),
                       "counts": res[:, :12].tolist(), "dsigma": res[:, 12:].tolist()}, f)
    if a.stream > 0:
        idx = np.arange(0, a.stream, a.stride)
        pts = prior_points(a.stream)[idx]
        res = batch_golden(pts, a.threads)
        with open(os.path.join(HERE, "golden_stream.json"), "w") as f:
            json.dump({"seed": 0x59334331, "stream_points": a.stream, "stride": a.stride,
                       "index": idx.tolist(), "points": pts.tolist(),
                       "values": res.tolist()}, f)


def batch_golden(pts, threads):
    """The 192 golden values (12 counts, then 180 Delta-Sigma) of each point."""
    plist = []
    for om, amp, slope, sln, fmis in pts:
        plist.append(O.default_params(omega_m=om, hmf_amp=amp, mor_slope=slope,
                                      mor_sigma_ln=sln, f_mis=fmis))
    arr = (O.Params * len(plist))(*plist)
    spec = O.spec_wide()
    lo = np.ascontiguousarray(LO_EDGES)
    zo = np.ascontiguousarray(ZO_EDGES)
    rr = np.ascontiguousarray(RADII)
    stride = 4 * 3 * (1 + len(RADII))
    res = np.zeros(len(plist) * stride)
    t1 = time.time()
    rc = O.lib.y3o_batch_golden(arr, C.byref(spec), len(plist), O.COUNT5D, O.ptr(lo), 4,
                                O.ptr(zo), 3, O.ptr(rr), len(RADII), threads, O.ptr(res))
    assert rc == 0
    print(f"{len(plist)} prior points in {time.time() - t1:.1f}s", flush=True)
    return res.reshape(len(plist), stride)


def fiducial():
    out = {"lo_edges": LO_EDGES, "zo_edges": ZO_EDGES, "radii": RADII}
    t0 = time.time()
    p0 = O.default_params(f_prj=0.0)
    t = O.OracleTables(p0, O.spec_config0())
    out["config0"] = {"params": params_dict(p0),
                      "value": float(O.golden(p0, t, O.COUNT3D, [20.0, 30.0])[0][0, 0])}
    p = O.default_params()
    tw = O.OracleTables(p, O.spec_wide())
    counts, ds = O.golden(p, tw, O.COUNT5D, LO_EDGES, ZO_EDGES, RADII)
    out["config1"] = {"params": params_dict(p), "values": counts.tolist()}
    out["config2"] = {"params": params_dict(p), "values": ds.tolist()}
    with open(os.path.join(HERE, "golden_fiducial.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"fiducial golden values in {time.time() - t0:.1f}s", flush=True)


if __name__ == "__main__":
    main()
