"""Autotuner of the batched cubature for the device it runs on.

    python -m y3_cluster_cpp_amd.autotune [--points 1:16,17:32] [--out PATH]

It tunes on parameter points of the prior-box stream that the benchmark
does not check: by default the 60 of points 0-63 that are not a multiple
of 16 (the benchmark checks every 16th point, 0, 16, 32, 48, ...; it times
all of them), 11,520 integrals. Their FP64 reference
values come from `tests/golden/golden_points.json`. It chooses:

1. **Region budgets by tolerance.**
   - Every calibration integral is run to each candidate region budget
     (4096 ... 65536) with its tolerance out of reach, so that it uses the
     whole budget.
   - A budget B* for every tolerance <= 1e-4: the smallest budget that
     meets two conditions, over the calibration integrals with their
     eps_rel capped at 1e-4:
     - it keeps at least 97% of the integrals that the largest budget
       gets within tolerance;
     - every reported error bounds the true error.
   - Bands inside it (eps_rel <= 2e-5, 5e-5, 1e-4, capped) and looser
     bands (eps_rel <= 2e-4, 5e-4, 1e-3) each get the smallest budget
     <= B* that meets the same two conditions on their own integrals.
   - This keeps accuracy. Budgets chosen for throughput alone would leave
     many more integrals outside their tolerance.
2. **Budget extension.** The tight bands' budgets (eps_rel <= 1e-4) are
   scaled by f (0.25 ... 1),
   and an integral whose reported error is still above ratio x eps_rel
   when its budget would cut its next step gets the budget doubled, as
   often as that holds, up to ext regions (65536 or 131072;
   y3g_options.ext_regions / ext_ratio).  Each setting runs the
   calibration integrals at their capped tolerances.  Of the settings that
   count at least 99.28% of them (99% and a margin for the sample's size)
   with no reported error below the true error, the one with the most
   counted integrals per unit of work (evaluations x FLOP per evaluation)
   is kept.  Most integrals meet their tolerance at a small budget while
   |R7 - R5| still over-states their error; the ones the base budget
   leaves far from it are the ones that need the larger budget.
3. **Launch shape.** The workgroup size (256 or 512 threads) and the
   regions split per step (256, 512 or 1024), chosen by counted integrals
   per second under those budgets. Shapes with any error that does not
   bound the true error are rejected.

The result is a JSON file read by `bench.py` and by `load()`. Its options
map onto `y3g_options` (`n_bands`, `band_eps`, `band_regions`,
`block_threads`, `max_split_per_step`, `max_regions`, `ext_regions`,
`ext_ratio`).
"""
import argparse
import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
DEFAULT_PATH = os.path.join(HERE, "tuning", "gfx950.json")
BUDGETS = (4096, 8192, 12288, 16384, 24576, 32768, 65536)
TIGHT_EDGES = (2e-5, 5e-5, 1e-4)
LOOSE_EDGES = (2e-4, 5e-4, 1e-3)
KEEP = 0.97  # share of the largest budget's in-tolerance integrals kept
# largest extended budgets tried.  262144 is left out: its work per counted
# integral looked lower on the calibration points, but on the benchmark's
# stream (up to 262144 regions, tight bands x 0.75, ratio 10) it ran 42.3 s
# at 99.17% against 40.0 s at 99.30% for up to 131072 at full bands -- a
# 31.7 MB region slab per workgroup makes every evaluation of the long
# integrals slower, which the work count does not see
EXT_REGIONS = (65536, 131072)
BAND_SCALES = (0.25, 0.375, 0.5, 0.75, 1.0)  # base budgets: the bands' x f
EXT_RATIOS = (3.0, 5.0, 10.0, 20.0)    # reported error / eps that extends
# counted share of the capped calibration integrals sought: 99% plus a
# margin of 3 standard errors of a 99% share over the calibration's n
# integrals (0.9928 for the default 60 points, 11,520 integrals), so that
# the benchmark's own sample of the stream stays above 99% (a setting at
# 99.17% here counted 98.97% there)
def target(n):
    return 0.99 + 3.0 * (0.99 * 0.01 / n) ** 0.5
# algorithmic FP32 FLOP per evaluation of the 5-D count and the 6-D
# Delta-Sigma integrand (the benchmark's tallies), as the work measure
FLOPS_PER_EVAL = {1: 83.0, 2: 136.0}


def load(path=DEFAULT_PATH):
    """The tuned options as y3g_options keyword arguments, or None."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    out = {"bands": [tuple(b) for b in t["bands"]], "block_threads": t["block_threads"],
           "max_split_per_step": t["max_split_per_step"], "max_regions": t["max_regions"]}
    if t.get("ext_regions", 0) > 0:
        out["ext_regions"] = int(t["ext_regions"])
        out["ext_ratio"] = float(t["ext_ratio"])
        out["max_regions"] = max(out["max_regions"], out["ext_regions"])
    return out


def _calibration(y3, ks):
    from y3_cluster_cpp_amd import workloads as W
    with open(os.path.join(ROOT, "tests", "golden", "golden_points.json")) as f:
        g = json.load(f)
    pts = np.array(g["points"])
    ref = np.concatenate([np.array(g["counts"]), np.array(g["dsigma"])], axis=1)
    eps = W.eps_stream(len(pts))
    spec = y3.spec_wide()
    params = [W.point_params(y3, pts[k]) for k in ks]
    igs = []
    for local, k in enumerate(ks):
        igs += W.point_integrals(y3, params[local], spec, local, eps[k], y3.PAGANI)
    return params, spec, igs, ref[ks].reshape(-1), eps[ks].reshape(-1)


def default_points():
    """The golden points the benchmark does not check: points 0-63 of the
    stream except every 16th (0, 16, 32, 48), 60 points."""
    return [k for k in range(64) if k % 16]


def tune(ks=None, out=DEFAULT_PATH, log=print):
    import y3_cluster_cpp_amd as y3
    rt = y3.Runtime(0)
    ks = list(ks) if ks is not None else default_points()
    params, spec, igs, ref, eps = _calibration(y3, ks)
    rt.tables_build(params, spec)
    n = len(igs)
    # 1. budgets: every integral run to each budget (tolerance out of reach)
    forced = []
    for g in igs:
        h = y3.Integral.from_buffer_copy(g)
        h.eps_rel = 1e-14
        forced.append(h)
    arr = (y3.Integral * n)(*forced)
    rel, honest, secs = {}, {}, {}
    rt.integrate(arr, y3.options(max_regions=BUDGETS[0]))  # first-launch costs
    for b in BUDGETS:
        t0 = time.perf_counter()
        r = rt.integrate(arr, y3.options(max_regions=b))
        secs[b] = time.perf_counter() - t0
        err = np.abs(r["value"] - ref)
        rel[b] = err / np.abs(ref)
        honest[b] = r["error"] >= err
        log(f"budget {b:6d}: {secs[b]:.3f} s, within eps {int((rel[b] <= eps).sum())}/{n}, "
            f"dishonest {int((~honest[b]).sum())}")
    def pick(mask, tol, limit):
        ok = {b: ((rel[b] <= tol) & honest[b])[mask] for b in BUDGETS}
        frac = {b: float(ok[b].mean()) for b in BUDGETS}
        best = max(frac[b] for b in BUDGETS if b <= limit)
        for b in BUDGETS:
            if b <= limit and frac[b] >= KEEP * best and honest[b][mask].all():
                return b, frac
        return limit, frac

    tight = np.ones(n, dtype=bool)
    bstar, frac = pick(tight, np.minimum(eps, 1e-4), BUDGETS[-1])
    table = [{"eps_le": 1e-4, "integrals": n, "budget": bstar,
              "note": "all integrals, eps_rel capped at 1e-4",
              "counted_frac": {str(b): frac[b] for b in BUDGETS}}]
    log(f"tolerances <= 1e-4: budget {bstar}")
    # bands inside it, on the integrals' capped tolerances (an integral whose
    # eps_rel exceeds 1e-4 falls in the last one when capped)
    bands = []
    capped = np.minimum(eps, 1e-4)
    lo = 0.0
    for edge in TIGHT_EDGES:
        m = (capped > lo) & (capped <= edge)
        lo = edge
        if not m.any():
            continue
        b, frac = pick(m, capped, bstar)
        bands.append((edge, b))
        table.append({"eps_le": edge, "integrals": int(m.sum()), "budget": b,
                      "counted_frac": {str(k): frac[k] for k in BUDGETS}})
        log(f"band eps <= {edge:.1e}: {int(m.sum())} integrals, budget {b}")
    lo = 1e-4
    for edge in LOOSE_EDGES:
        m = (eps > lo) & (eps <= edge)
        lo = edge
        if not m.any():
            continue
        b, frac = pick(m, eps, bstar)
        bands.append((edge, b))
        table.append({"eps_le": edge, "integrals": int(m.sum()), "budget": b,
                      "counted_frac": {str(k): frac[k] for k in BUDGETS}})
        log(f"band eps <= {edge:.1e}: {int(m.sum())} integrals, budget {b}")
    max_regions = max(b for _, b in bands)
    # 2. budget extension: the bands' budgets scaled by f, and integrals whose
    # reported error is still above ratio x eps when their budget is spent
    # continue to ext regions; the fastest setting that counts at least
    # target(n) of the capped calibration integrals with no reported error below
    # the true error (else the one that counts the most)
    capped_arr = []
    for g in igs:
        h = y3.Integral.from_buffer_copy(g)
        h.eps_rel = min(h.eps_rel, 1e-4)
        capped_arr.append(h)
    capped_arr = (y3.Integral * n)(*capped_arr)
    flops = np.array([FLOPS_PER_EVAL.get(g.kind, 0) for g in igs], dtype=np.float64)
    ext_runs = []
    for x in EXT_REGIONS:
        # first launch at this slab size (allocation), untimed
        rt.integrate(capped_arr, y3.options(bands=bands, max_regions=x, ext_regions=x,
                                            ext_ratio=EXT_RATIOS[-1]))
        for f in BAND_SCALES:
            # the tight bands' budgets scale; the looser ones keep theirs
            # (chosen above with every reported error bounding its true one)
            bf = [(e, max(BUDGETS[0], int(round(b * f / 2048)) * 2048) if e <= 1e-4 else b)
                  for e, b in bands]
            for ratio in EXT_RATIOS:
                o = y3.options(bands=bf, max_regions=x, ext_regions=x, ext_ratio=ratio)
                t0 = time.perf_counter()
                r = rt.integrate(capped_arr, o)
                dt = time.perf_counter() - t0
                err = np.abs(r["value"] - ref)
                ok = (err <= capped * np.abs(ref)) & (r["error"] >= err)
                dis = int((r["error"] < err).sum())
                # work: evaluations x the kind's FLOP per evaluation (a
                # stream step balances its integrals over the workgroups,
                # so its time follows the work; these few integrals' launch
                # time follows the slowest one instead)
                work = float((r["neval"] * flops).sum())
                ext_runs.append({"ext_regions": x, "band_scale": f, "ext_ratio": ratio,
                                 "bands": bf, "seconds": dt, "counted": int(ok.sum()),
                                 "counted_frac": float(ok.mean()), "gflop": work / 1e9,
                                 "counted_per_tflop": float(ok.sum() / work * 1e12),
                                 "dishonest": dis})
                log(f"ext {x} scale {f} ratio {ratio}: {dt:.3f} s, {work / 1e12:.2f} TFLOP, "
                    f"counted {int(ok.sum())}/{n}, dishonest {dis}")
    honest_runs = [e for e in ext_runs if e["dishonest"] == 0] or ext_runs
    meets = [e for e in honest_runs if e["counted_frac"] >= target(n)]
    ext = (max(meets, key=lambda e: e["counted_per_tflop"]) if meets
           else max(honest_runs, key=lambda e: (e["counted"], e["counted_per_tflop"])))
    bands = [tuple(b) for b in ext["bands"]]
    max_regions = ext["ext_regions"]
    log(f"extension: bands x {ext['band_scale']}, ratio {ext['ext_ratio']}, "
        f"{ext['ext_regions']} regions: {ext['counted_frac']:.4f} counted")
    # 3. launch shape under those budgets, with the integrals' own tolerances
    arr = (y3.Integral * n)(*igs)
    shapes = []
    for block in (256, 512):
        for kmax in (256, 512, 1024):
            o = y3.options(bands=bands, max_regions=max_regions, block_threads=block,
                           max_split_per_step=kmax, ext_regions=ext["ext_regions"],
                           ext_ratio=ext["ext_ratio"])
            rt.integrate(arr, o)
            t0 = time.perf_counter()
            r = rt.integrate(arr, o)
            dt = time.perf_counter() - t0
            err = np.abs(r["value"] - ref)
            ok = (err <= np.minimum(eps, 1e-4) * np.abs(ref)) & (r["error"] >= err)
            dis = int((r["error"] < err).sum())
            shapes.append({"block_threads": block, "max_split_per_step": kmax, "seconds": dt,
                           "counted": int(ok.sum()), "counted_per_s": float(ok.sum() / dt),
                           "dishonest": dis})
            log(f"block {block} kmax {kmax}: {dt:.3f} s, counted {int(ok.sum())}, "
                f"dishonest {dis}")
    good = [s for s in shapes if s["dishonest"] == 0] or shapes
    best = max(good, key=lambda s: s["counted_per_s"])
    info = rt.device_info()
    rt.close()
    res = {"device": info["name"], "calibration_points": ks,
           "bands": [[e, b] for e, b in bands], "max_regions": max_regions,
           "ext_regions": ext["ext_regions"], "ext_ratio": ext["ext_ratio"],
           "extension_target": target(n), "extension_sweep": ext_runs,
           "block_threads": best["block_threads"],
           "max_split_per_step": best["max_split_per_step"],
           "budget_sweep": {str(b): {"seconds": secs[b],
                                     "within_eps": int((rel[b] <= eps).sum()),
                                     "dishonest": int((~honest[b]).sum())} for b in BUDGETS},
           "bands_detail": table, "shapes": shapes}
    os.makedirs(os.path.dirname(out), exist_ok=True)
    with open(out, "w") as f:
        json.dump(res, f, indent=1)
    log(f"wrote {out}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--points", default="",
                    help="calibration points as ranges, e.g. 1:16,17:32 (default: the 60 "
                         "golden points the benchmark does not check)")
    ap.add_argument("--out", default=DEFAULT_PATH)
    a = ap.parse_args()
    ks = None
    if a.points:
        ks = []
        for part in a.points.split(","):
            lo, hi = (int(v) for v in part.split(":"))
            ks += list(range(lo, hi))
    tune(ks, a.out)


if __name__ == "__main__":
    main()
