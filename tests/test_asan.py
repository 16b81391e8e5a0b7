"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (no GPU).

* the C oracle, driven on every problem kind, the TSPJ local search and the
  NSGA-II diversity injection, through tests/cpp/oracle_asan_test.c with
  buffers of exactly the documented sizes;
* the CPU baseline loop (oracle/cpu_reference_loop.cpp) on a small instance
  with islands and migration;
* the facade's file loaders (tests/cpp/loader_test.cpp) on the fixture files,
  whose output must equal the optimised build's.
Each program is built with g++/gcc only (no device code) and must exit 0
with no sanitizer report.
"""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
ENGINE = os.path.join(ROOT, "llm-brkga_amd")
ENV = dict(os.environ, ASAN_OPTIONS="abort_on_error=0:halt_on_error=1",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def _clean(res):
    assert res.returncode == 0, res.stderr[-4000:]
    assert "AddressSanitizer" not in res.stderr and "runtime error" not in res.stderr, res.stderr[-4000:]


@pytest.fixture(scope="module")
def oracle_bins():
    subprocess.run(["make", "-C", ORACLE, "asan"], check=True, capture_output=True, timeout=600)
    return os.path.join(ORACLE, "_asan")


def test_oracle_every_kind(oracle_bins):
    res = subprocess.run([os.path.join(oracle_bins, "oracle_asan_test")], capture_output=True, text=True,
                         timeout=300, env=ENV)
    _clean(res)
    lines = res.stdout.splitlines()
    assert lines[-1] == "done"
    # NaN keys order last, ties keep index order
    assert lines[0] == "argsort 7 4 2 8 0 3 6 1 5"
    names = {l.split()[0] for l in lines[1:-1]}
    for k in ("tsp_coords", "tsp_matrix", "knapsack_threshold", "knapsack_greedy", "tspj", "bitsp_coords",
              "bitsp_matrix", "bitsp_three_cities", "sphere", "rastrigin", "rosenbrock", "jobshop", "zdt1",
              "vrp_days", "movrp"):
        assert k in names
    three = [l for l in lines if l.startswith("bitsp_three_cities")][0].split()
    assert int(three[-1]) == 5  # every step of a 3-city instance injects


def test_cpu_reference_loop(oracle_bins, tmp_path):
    xy = np.random.default_rng(1).uniform(0, 100, (40, 2)).astype(np.float32)
    f = tmp_path / "xy.bin"
    xy.tofile(f)
    res = subprocess.run([os.path.join(oracle_bins, "cpu_reference_loop"), str(f), "40", "32", "3", "6", "5", "0.7",
                          "6", "2", "2", "1", "2", "7"], capture_output=True, text=True, timeout=300, env=ENV)
    _clean(res)
    out = json.loads(res.stdout)
    assert out["generations"] == 6 and len(out["runs_s"]) == 2 and out["best"] > 0


def test_loaders(tmp_path):
    plain = os.path.join(ENGINE, "_build", "loader_test")
    asan = plain + "_asan"
    subprocess.run(["make", "-C", ENGINE, "_build/loader_test", "_build/loader_test_asan"], check=True,
                   capture_output=True, timeout=900)
    keys = np.load(os.path.join(ROOT, "tests", "golden", "loaders_ref.npz"))["tiny8_keys"]
    kf = tmp_path / "keys.txt"
    np.savetxt(kf, keys, fmt="%.9g")
    data = os.path.join(ROOT, "tests", "golden", "data")
    res = subprocess.run([asan, data, str(kf)], capture_output=True, text=True, timeout=300, env=ENV)
    _clean(res)
    ref = subprocess.run([plain, data, str(kf)], check=True, capture_output=True, text=True, timeout=60).stdout
    assert json.loads(res.stdout) == json.loads(ref)
