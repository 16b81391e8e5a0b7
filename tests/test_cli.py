"""Prebuilt command-line solver (llm-brkga_amd/brkga_amd/brkga_cli): the
flags and stdout lines of the reference's driver (main.cu:46-305) and of
the mains generated by llm_solver/core/execution_manager.py, whose regexes
read "Best fitness:", "Generations:" and "Pareto front size:"."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "llm-brkga_amd", "brkga_amd", "brkga_cli")
DATA = os.path.join(ROOT, "tests", "golden", "data")


@pytest.fixture(scope="module")
def cli():
    if not os.path.exists(CLI):
        subprocess.run(["make", "-j8", "-C", os.path.join(ROOT, "llm-brkga_amd"), "brkga_amd/brkga_cli"], check=True,
                       capture_output=True)
    return CLI


def run(cli, args, cwd=None, env=None, timeout=120):
    return subprocess.run([cli] + args, capture_output=True, text=True, cwd=cwd, env=env, timeout=timeout)


def test_usage_and_argument_errors(cli):
    r = run(cli, ["--help"])
    assert r.returncode == 0 and "--data <file>" in r.stdout and "multi_tsp_config" in r.stdout
    r = run(cli, [])
    assert r.returncode == 1 and "Missing required arguments" in r.stderr
    r = run(cli, ["--data", "/nonexistent.tsp", "--config", "tsp_config"])
    assert r.returncode == 1 and "Data file not found" in r.stderr
    r = run(cli, ["--data", os.path.join(DATA, "tiny8.tsp"), "--config", "vrp_config"])
    assert r.returncode == 1 and "Unknown config type" in r.stderr


CASES = [("tsp_config", "tiny8.tsp"), ("knapsack_config", "knap_small.txt"), ("knapsack_config", "knap_cap.csv"),
         ("tspj_config", "mini_TT.csv"), ("multi_tsp_config", "tiny8.tsp"),
         ("multi_objective_vrp_config", "tiny-n12-k3.vrp")]


@pytest.mark.gpu
@pytest.mark.parametrize("config,data", CASES)
def test_cli_solves(cli, tmp_path, config, data):
    print(f"cli {config} {data}", flush=True)
    r = run(cli, ["--data", os.path.join(DATA, data), "--config", config, "--pop", "64", "--gen", "20",
                  "--seed", "7"], cwd=tmp_path)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    out = r.stdout
    assert "Optimization completed successfully" in out
    assert int(re.search(r"Generations:\s*(\d+)", out).group(1)) == 20
    if config in ("multi_tsp_config", "multi_objective_vrp_config"):
        assert int(re.search(r"Pareto front size:\s*(\d+)", out).group(1)) >= 1
        front = (tmp_path / f"{config}_pareto_front.txt").read_text()
        assert "# Pareto Front" in front
        if config == "multi_objective_vrp_config":
            routes = (tmp_path / "pareto_solutions" / "routes_data.txt").read_text()
            assert "SOLUTION 0" in routes and "END_ROUTE" in routes
            assert (tmp_path / "pareto_solutions" / "solution_000.txt").exists()
    else:
        assert re.search(r"Best fitness:\s*([-\d.]+)", out)
        assert (tmp_path / "solutions" / f"{config}_solution.txt").exists()


@pytest.mark.gpu
def test_cli_quick_test_env(cli, tmp_path):
    print("cli QUICK_TEST", flush=True)
    env = dict(os.environ, QUICK_TEST="3")
    r = run(cli, ["--data", os.path.join(DATA, "tiny8.tsp"), "--config", "tsp_config", "--pop", "32"], cwd=tmp_path,
            env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Quick test mode: limiting to 3 generations" in r.stdout
    assert int(re.search(r"Generations:\s*(\d+)", r.stdout).group(1)) == 3
