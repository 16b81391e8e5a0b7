"""Shared test setup: the `gpu` marker, import paths and instance helpers.

Tests marked `gpu` need a real device and call the engine through its C ABI
(llm-brkga_amd/brkga_amd/libbrkga_hip.so); everything else runs on the CPU.
The oracle (oracle/) is the checker every GPU result is compared with.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "llm-brkga_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine through the C ABI)")


@pytest.fixture(scope="session")
def engine():
    import brkga_amd

    if brkga_amd.device_count() == 0:
        pytest.fail("no GPU visible to the HIP engine")
    eng = brkga_amd.Engine(0)
    yield eng


def rng(seed):
    return np.random.default_rng(seed)


def random_keys(seed, rows, cols, ties=False):
    """Engine-style 24-bit keys; with ties=True many equal keys per row."""
    g = rng(seed)
    if ties:
        return (g.integers(0, 8, size=(rows, cols)).astype(np.float32) / np.float32(8.0)).astype(np.float32)
    return (g.integers(0, 1 << 24, size=(rows, cols)).astype(np.float32) * np.float32(2.0 ** -24)).astype(np.float32)
