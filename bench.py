"""Throughput of the cluster-observable integration path on one MI355X.

Workload: the batched prior-box stream -- parameter points drawn from the
box Omega_m in [0.2, 0.4], HMF amplitude in [0.7, 0.9], MOR slope in
[0.7, 1.1], sigma_ln in [0.1, 0.4], f_mis in [0, 0.4] (Philox key
0x59334331), each point carrying 192 integrals: 12 five-dimensional number
counts (4 lambda_obs x 3 z_obs bins) and 180 six-dimensional miscentred
Delta-Sigma integrals (12 bins x 15 radii log-spaced in [0.1, 30] Mpc/h),
per-integral eps_rel log-uniform in [1e-5, 1e-3], capped at 1e-4 for the
headline.  Per-point tables (HMF 200x32, D_A 512, Sigma_crit^-1 128) are
rebuilt on the device inside each step.

A step is one batch of parameter points taken in stream order.  An integral
is counted when |I - I_oracle| <= eps_rel |I_oracle|, that error is <= 1e-4,
and the reported error bounds it; I_oracle are the FP64 CPU oracle's golden
values of the stream's first points (tests/golden/golden_points.json), so
every integral of the timed steps is checked.

    python bench.py [--steps K] [--warmup W] [--points-per-step B]
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

LO_EDGES = [20.0, 30.0, 45.0, 60.0, 300.0]
ZO_EDGES = [0.20, 0.35, 0.50, 0.65]
RADII = list(np.logspace(np.log10(0.1), np.log10(30.0), 15))
SEED = 0x59334331

# FP32 operations per integrand evaluation of the device integrands
# (+ - * / = 1, FMA = 2) and transcendental (SFU) calls, counted from
# csrc/y3g_device.hpp integrand<float>.
FLOPS_PER_EVAL = {0: 55, 1: 80, 2: 165}
SFU_PER_EVAL = {0: 7, 1: 9, 2: 24}
PEAK_FP32 = 157.3e12  # MI355X dense FP32 vector, FLOP/s
# HBM bytes per integrand evaluation of k_cubature<float,512>: rocprofv3
# FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 (factors from a 256 MiB copy) over
# the evaluations of the same launch, profiles/k_cubature_prior_stream_pmc.md
# (7.08e9 B / 1.829e9 evaluations).  Region-slab traffic; the integrand reads
# only LDS-resident tables.
HBM_BYTES_PER_EVAL = 7.0752e9 / 1.829061504e9


def prior_points(n, seed=SEED):
    rng = np.random.Generator(np.random.Philox(key=seed))
    u = rng.random((n, 5))
    lo = np.array([0.2, 0.7, 0.7, 0.1, 0.0])
    hi = np.array([0.4, 0.9, 1.1, 0.4, 0.4])
    return lo + u * (hi - lo)


def eps_stream(n_points, seed=SEED):
    """Per-integral eps_rel, log-uniform in [1e-5, 1e-3], 192 per point."""
    rng = np.random.Generator(np.random.Philox(key=seed + 1))
    return 10.0 ** rng.uniform(-5.0, -3.0, (n_points, 192))


def point_integrals(y3, params, spec, k, eps_row, algorithm):
    igs = []
    n = 0
    for i in range(4):
        for j in range(3):
            g = y3.make_integral(params, spec, y3.COUNT5D, LO_EDGES[i], LO_EDGES[i + 1],
                                 ZO_EDGES[j], ZO_EDGES[j + 1], point=k,
                                 eps_rel=float(eps_row[n]), algorithm=algorithm)
            igs.append(g)
            n += 1
    for i in range(4):
        for j in range(3):
            for r in RADII:
                g = y3.make_integral(params, spec, y3.DSIGMA6D, LO_EDGES[i], LO_EDGES[i + 1],
                                     ZO_EDGES[j], ZO_EDGES[j + 1], r, point=k,
                                     eps_rel=float(eps_row[n]), algorithm=algorithm)
                igs.append(g)
                n += 1
    return igs


def load_golden():
    path = os.path.join(ROOT, "tests", "golden", "golden_points.json")
    with open(path) as f:
        g = json.load(f)
    pts = np.array(g["points"])
    ref = np.concatenate([np.array(g["counts"]), np.array(g["dsigma"])], axis=1)
    return pts, ref


def count_ok(res, ref, eps):
    err = np.abs(res["value"] - ref)
    rel = err / np.abs(ref)
    ok = (rel <= eps) & (rel <= 1e-4) & (res["error"] >= err)
    return ok, rel


def cpu_baseline(points, eps, n_sample, threads):
    """The oracle's FP64 integrands under its CPU restatement of the same
    cubature, one integral per thread (OpenMP), on a bounded sample."""
    import oracle as O
    k = 0
    om, amp, slope, sln, fmis = points[k]
    op = O.default_params(omega_m=om, hmf_amp=amp, mor_slope=slope, mor_sigma_ln=sln,
                          f_mis=fmis)
    spec = O.spec_wide()
    igs = []
    epsl = []
    # sample: the 12 count integrals and the first Delta-Sigma radii, in order
    n = 0
    for i in range(4):
        for j in range(3):
            igs.append(O.make_integral(op, spec, O.COUNT5D, LO_EDGES[i], LO_EDGES[i + 1],
                                       ZO_EDGES[j], ZO_EDGES[j + 1], 1.0))
            epsl.append(eps[k, n])
            n += 1
    for i in range(4):
        for j in range(3):
            for r in RADII:
                igs.append(O.make_integral(op, spec, O.DSIGMA6D, LO_EDGES[i], LO_EDGES[i + 1],
                                           ZO_EDGES[j], ZO_EDGES[j + 1], r))
                epsl.append(eps[k, n])
                n += 1
    cnt = list(range(0, 12, 2))
    dsg = list(range(12, 192, 11))
    idx = [v for pair in zip(cnt, dsg) for v in pair][:n_sample]
    arr = (O.Integral * len(idx))(*[igs[i] for i in idx])
    pt = (C.c_int * len(idx))(*([0] * len(idx)))
    e = np.ascontiguousarray([epsl[i] for i in idx])
    out = (O.Result * len(idx))()
    pa = (O.Params * 1)(op)
    sa = (O.TableSpec * 1)(spec)
    t0 = time.perf_counter()
    used = O.lib.y3o_batch_integrate(pa, sa, 1, arr, pt, O.ptr(e), len(idx), 0,
                                     C.byref(O.cub_opts(max_regions=262144, max_eval=1 << 40)),
                                     C.byref(O.vegas_opts()), threads, out)
    dt = time.perf_counter() - t0
    return len(idx) / dt, used, dt, len(idx)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--points-per-step", type=int, default=16)
    ap.add_argument("--eps-cap", type=float, default=1e-4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-sample", type=int, default=6)
    ap.add_argument("--cpu-sample", type=int, default=6)
    ap.add_argument("--max-regions", type=int, default=0,
                    help="per-integral region budget (0: library default)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 256 MiB device copy (calibrates FETCH_SIZE / "
                         "WRITE_SIZE counters) and exit")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch
        n = 256 << 20
        src = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n}))
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every box here has one MI355X and the "
              "integrals shard by parameter point with no exchange (see DESIGN.md); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import y3_cluster_cpp_amd as y3
    pts, ref = load_golden()
    B = a.points_per_step
    need = B * max(a.steps, 1)
    if need > len(pts):
        raise SystemExit(f"{need} points needed, {len(pts)} have oracle values")
    eps = np.minimum(eps_stream(len(pts)), a.eps_cap)
    spec = y3.spec_wide()
    params_all = [y3.default_params(omega_m=om, hmf_amp=amp, mor_slope=sl, mor_sigma_ln=sn,
                                    f_mis=fm) for om, amp, sl, sn, fm in pts]
    rt = y3.Runtime(0)
    info = rt.device_info()
    opts = y3.options(precision=y3.FP32SPLIT)
    if a.max_regions > 0:
        opts.max_regions = a.max_regions

    def batch(step):
        ks = range(step * B, (step + 1) * B)
        igs = []
        for local, k in enumerate(ks):
            igs += point_integrals(y3, params_all[k], spec, local, eps[k], y3.PAGANI)
        return list(ks), (y3.Integral * len(igs))(*igs)

    batches = [batch(s) for s in range(a.steps)]
    # warmup (untimed): the first batch
    for _ in range(a.warmup):
        ks, arr = batches[0]
        rt.tables_build([params_all[k] for k in ks], spec)
        rt.integrate(arr, opts)

    results = []
    kern_ms = 0.0
    flops = 0.0
    nev_total = 0
    t0 = time.perf_counter()
    for s in range(a.steps):
        ks, arr = batches[s]
        rt.tables_build([params_all[k] for k in ks], spec)
        r = rt.integrate(arr, opts)
        ms, nev = rt.last_timing()
        kern_ms += ms
        nev_total += nev
        kinds = np.array([arr[i].kind for i in range(len(arr))])
        flops += float(sum(r["neval"][kinds == k].sum() * FLOPS_PER_EVAL[k] for k in (1, 2)))
        results.append(r)
    elapsed = time.perf_counter() - t0

    ok_all = []
    rel_all = []
    for s in range(a.steps):
        ks = list(range(s * B, (s + 1) * B))
        ok, rel = count_ok(results[s], ref[ks].reshape(-1), eps[ks].reshape(-1))
        ok_all.append(ok)
        rel_all.append(rel)
    ok_all = np.concatenate(ok_all)
    rel_all = np.concatenate(rel_all)
    n_int = len(ok_all)
    counted = int(ok_all.sum())
    value = counted / elapsed

    # plain FP64 path: one integral per launch sequence, serial over integrals
    plain = None
    if not a.no_plain:
        ks, arr = batches[0]
        rt.tables_build([params_all[k] for k in ks], spec)
        # sample: alternate count and Delta-Sigma integrals of point 0
        cnt = list(range(0, 12, 2))
        dsg = list(range(12, 192, 11))
        idx = [v for pair in zip(cnt, dsg) for v in pair][: a.plain_sample]
        sub = (y3.Integral * len(idx))(*[arr[i] for i in idx])
        tp = time.perf_counter()
        popts = y3.options(precision=y3.FP64, single=1)
        popts.max_regions = opts.max_regions
        rp = rt.integrate(sub, popts)
        dtp = time.perf_counter() - tp
        okp, _ = count_ok(rp, ref[ks[0]][idx], eps[ks[0]][idx])
        plain = {"integrals_per_s": len(idx) / dtp, "counted_per_s": int(okp.sum()) / dtp,
                 "sample": f"{len(idx)} integrals of point 0 ({sum(1 for i in idx if i < 12)} "
                           f"counts, {sum(1 for i in idx if i >= 12)} Delta-Sigma)",
                 "seconds": dtp}

    cpu = None
    if not a.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        v, used, dt, n = cpu_baseline(pts, eps, a.cpu_sample, threads)
        cpu = {"value": v, "unit": "integrals/s", "cores": int(used), "kind": "port",
               "sample": f"{n} integrals of prior point 0 (counts and Delta-Sigma at their "
                         f"stream eps_rel), FP64 oracle + CPU Genz-Malik cubature, {dt:.1f} s"}

    achieved = flops / (kern_ms * 1e-3) if kern_ms > 0 else 0.0
    out = {
        "metric": "integrals/sec at rel err <=1e-4 vs FP64 oracle; p50/p99 request latency",
        "value": value,
        "unit": "integrals/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 integrand, fp64 accumulation",
        "data": "synthetic",
        "config": {"workload": "prior-box batched stream (192 integrals per parameter point: "
                               "12 5-D counts + 180 6-D miscentred Delta-Sigma)",
                   "points_per_step": B, "integrals_per_step": n_int // a.steps,
                   "eps_rel": "log-uniform [1e-5,1e-3] capped at %g" % a.eps_cap,
                   "integrator": "Genz-Malik cubature, workgroup per integral",
                   "counted": counted, "integrals": n_int,
                   "max_rel_err": float(rel_all.max()),
                   "device": info["name"]},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
                     "traffic": nev_total / a.steps * HBM_BYTES_PER_EVAL,
                     "kernel": "k_cubature<float,512>", "kernel_ms_per_step": kern_ms / a.steps,
                     "evals_per_s": nev_total / (kern_ms * 1e-3) if kern_ms > 0 else 0.0},
        "cpu_baseline": cpu,
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": (value / plain["counted_per_s"]) if plain and plain["counted_per_s"] > 0 else None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
