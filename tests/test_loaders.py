"""Problem-data loaders of the C++ facade (include/brkga/configs/*.hpp) on
the fixture files in tests/golden/data, no GPU involved.

Pins:
  * knapsack text / CSV loaders: equal to what the reference's own
    KnapsackConfig::load_from_file returns on the same files (compiled from
    configs/knapsack_config.hpp into oracle/_ref; vectors in
    tests/golden/loaders_ref.npz);
  * bi-objective TSP loader: the objectives of 24 tours, summed in the
    reference's order over the loaded distance and generated time matrices,
    equal those of the reference's own MultiTSPConfig::load_from_file
    (configs/multi_tsp_config.hpp:165-269), which pins both matrices and the
    std::mt19937(42) speed factors;
  * TSPLIB and TSPJ loaders (their reference headers need CUDA): the parsed
    values equal the numbers written in the fixture files.
"""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "llm-brkga_amd", "_build", "loader_test")
DATA = os.path.join(ROOT, "tests", "golden", "data")
GOLDEN = os.path.join(ROOT, "tests", "golden", "loaders_ref.npz")


@pytest.fixture(scope="module")
def loaded(tmp_path_factory):
    if not os.path.exists(BIN):
        subprocess.run(["make", "-j8", "-C", os.path.join(ROOT, "llm-brkga_amd"), "_build/loader_test"], check=True,
                       capture_output=True)
    keys = np.load(GOLDEN)["tiny8_keys"]
    kf = tmp_path_factory.mktemp("loaders") / "keys.txt"
    np.savetxt(kf, keys, fmt="%.9g")
    out = subprocess.run([BIN, DATA, str(kf)], check=True, capture_output=True, text=True, timeout=60).stdout
    return json.loads(out)


def test_tsplib_coordinates(loaded):
    assert loaded["tsp_name"] == "tiny8" and loaded["tsp_n"] == 8
    assert loaded["tsp_kind"] == 0  # coordinate instance, no n x n matrix
    # ids 1..8 in file order 1,3,2,4,5,6,(9 ignored),7,8
    expect = [565, 575, 25, 185, 345, 750, 945, 685, 845, 655, 880, 660, 25, 230, 525, 1000]
    assert loaded["tsp_xy"] == expect


@pytest.mark.parametrize("name", ["knap_small_txt", "knap_cap_csv", "knap_nocap_csv"])
def test_knapsack_matches_reference_loader(loaded, name):
    g = np.load(GOLDEN)
    assert np.array_equal(np.float32(loaded[name + "_w"]), g[name + "_w"])
    assert np.array_equal(np.float32(loaded[name + "_v"]), g[name + "_v"])
    assert np.array_equal(np.float32(loaded[name + "_cap"]), g[name + "_cap"])
    assert loaded[name + "_name"] == name.rsplit("_", 1)[0]


def test_tspj_pair_from_either_file(loaded):
    travel = [[0, 12.5, 30, 18], [12.5, 0, 21.25, 9], [30, 21.25, 0, 14], [18, 9, 14, 0]]
    jobs = [[0, 0, 0, 0], [0, 5, 7.5, 3], [0, 4, 6, 8], [0, 2, 9, 1.5]]
    for src in ("tt", "jt"):
        assert loaded[f"tspj_from_{src}_travel"] == travel
        assert loaded[f"tspj_from_{src}_jobs"] == jobs
        assert loaded[f"tspj_from_{src}_name"] == "mini" and loaded[f"tspj_from_{src}_n"] == 3


def test_multitsp_matches_reference_loader(loaded):
    assert loaded["multi_n"] == 7
    obj = np.float32(loaded["multi_objectives"]).reshape(-1, 2)
    assert np.array_equal(obj, np.load(GOLDEN)["tiny8_objectives"])
