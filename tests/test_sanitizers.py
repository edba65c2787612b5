"""Host code under AddressSanitizer, UndefinedBehaviorSanitizer and leak
detection (no GPU): the oracle's code paths (oracle/asan_check.c, built by
`make -C oracle asan`) and the CosmoSIS module layer's host test
(tests/cpp/test_module_host.cc: DataBlock readers, volumes, grid points,
integral lists, output layout, setup failing cleanly)."""
import os
import subprocess

from conftest import ROOT

ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1")


def test_oracle_under_sanitizers():
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "asan"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([os.path.join(ROOT, "oracle", "_build", "asan_check")],
                       capture_output=True, text=True, timeout=900, env=ENV)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "0 non-finite values" in r.stdout


def test_module_host_under_sanitizers(tmp_path):
    mod = os.path.join(ROOT, "y3_cluster_cpp_amd", "modules")
    lib = os.path.join(ROOT, "y3_cluster_cpp_amd")
    out = str(tmp_path / "test_module_host_san")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-Wall",
           "-I" + mod, "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "test_module_host.cc"), "-o", out,
           "-L" + mod, "-ldatablock", "-L" + lib, "-ly3g", "-ldl",
           "-Wl,-rpath," + mod, "-Wl,-rpath," + lib]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([out, os.path.join(mod, "SigmaMiscent6DIntegrand.so")],
                       capture_output=True, text=True, timeout=300, env=ENV)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert " 0 failures" in r.stdout
