"""The benchmark's step plan without a GPU: the driver's command line
(--gpus 1 --steps 20 --warmup 5) covers the 4096-point stream once, every
step samples the golden subset, and every descriptor builds."""
import json
import os
import subprocess
import sys

import numpy as np

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402
from y3_cluster_cpp_amd.workloads import N_PER_POINT, prior_points  # noqa: E402


def test_plan_covers_stream_once():
    gold, gold_pts = bench.load_stream_golden()
    checked = set(gold)
    timed, warm, checks = bench.plan_steps(4096, 20, 5, checked)
    allpts = np.concatenate(timed)
    assert np.array_equal(allpts, np.arange(4096))
    assert sorted(len(x) for x in timed) == [204] * 4 + [205] * 16
    assert len(warm) == 5 and all(w is timed[0] for w in warm)
    assert sum(len(c) for c in checks) == 256 and min(len(c) for c in checks) >= 12
    for sl, c in zip(timed, checks):
        assert set(c) <= set(int(k) for k in sl)


def test_golden_subset_is_the_stream():
    gold, gold_pts = bench.load_stream_golden()
    pts = prior_points(4096)
    assert sorted(gold) == list(range(0, 4096, 16))
    for k in gold:
        assert np.array_equal(pts[k], gold_pts[k])
        assert gold[k].shape == (N_PER_POINT,) and np.all(np.isfinite(gold[k]))
    # the first 64 points' file and the subset agree where they overlap
    p64, r64 = bench.load_golden()
    for k in (0, 16, 32, 48):
        assert np.array_equal(r64[k], gold[k])


def test_plan_steps_rejects_bad_counts():
    for steps in (0, 4097):
        try:
            bench.plan_steps(4096, steps, 0, set())
        except ValueError:
            continue
        raise AssertionError(steps)


def test_driver_command_plans_and_builds_descriptors():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "1",
                        "--steps", "20", "--warmup", "5", "--plan-only"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    plan = json.loads(r.stdout.strip().splitlines()[-1])
    assert plan["integrals"] == 786432 and plan["warmup"] == 5
    assert len(plan["steps"]) == 20 and sum(plan["steps"]) == 4096


def test_tuned_options_are_valid():
    """The committed tuning file loads into options the library accepts:
    ascending band edges, budgets within max_regions, an extension that the
    slab holds, and calibration points outside the benchmark's checked
    subset."""
    from y3_cluster_cpp_amd import autotune
    t = autotune.load()
    assert t is not None
    edges = [e for e, _ in t["bands"]]
    assert edges == sorted(edges) and len(edges) <= 8
    assert all(64 <= r <= t["max_regions"] for _, r in t["bands"])
    assert 0 < t["ext_regions"] <= t["max_regions"] and t["ext_ratio"] > 0
    assert t["block_threads"] in (256, 512) and 1 <= t["max_split_per_step"] <= 1024
    with open(autotune.DEFAULT_PATH) as f:
        raw = json.load(f)
    assert not any(k % 16 == 0 for k in raw["calibration_points"])
    assert raw["extension_target"] > 0.99
