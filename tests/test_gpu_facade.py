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


def test_cpp_facade():
    print("C++ facade program", flush=True)
    assert os.path.exists(BIN), f"facade test program not built: {BIN} (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout[-4000:], r.stderr[-2000:])
    assert r.returncode == 0 and "FACADE OK" in r.stdout
