"""C++ facade (include/brkga/): Solver / Population / configs driven as a
caller of the reference's API would, on one GPU.  The compiled program
(tests/cpp/facade_test.cpp, built next to the engine library) checks the
facade against the C ABI it sits on: same best fitness for the same seed,
callbacks equal to the device results, valid decodes, a non-dominated
Pareto front."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "llm-brkga_amd", "_build", "facade_test")
HOST_BIN = os.path.join(ROOT, "llm-brkga_amd", "_build", "host_config_test")


def test_cpp_facade():
    print("C++ facade program", flush=True)
    assert os.path.exists(BIN), f"facade test program not built: {BIN} (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout[-4000:], r.stderr[-2000:])
    assert r.returncode == 0 and "FACADE OK" in r.stdout


def test_cpp_host_evaluated_configs():
    """Configs in the reference generator's template form (host lambdas, no
    engine decoder) compile unchanged and run; a host greedy-knapsack config
    follows the device greedy decoder's trajectory bit for bit."""
    print("C++ host-evaluated configs", flush=True)
    assert os.path.exists(HOST_BIN), f"host config test program not built: {HOST_BIN}"
    r = subprocess.run([HOST_BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout[-4000:], r.stderr[-2000:])
    assert r.returncode == 0 and "HOST CONFIGS OK" in r.stdout
