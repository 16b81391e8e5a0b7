"""Pins the oracle's Philox4x32-10 stream and key conversion.

Checks: the published Random123 known-answer vectors (tests/golden/philox_kat.json),
AMD's open-source rocRAND Philox engine (compiled host-only from
/opt/rocm/include/rocrand), and the package's own numpy Philox used to draw
synthetic instances.  The reference itself uses cuRAND XORWOW, which is closed
and not reproduced (core/cuda_kernels.cuh:10-15).
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import pyoracle
from brkga_amd import workloads

HERE = os.path.dirname(os.path.abspath(__file__))


def test_philox_known_answers():
    with open(os.path.join(HERE, "golden", "philox_kat.json")) as f:
        kats = json.load(f)
    assert len(kats) == 3
    for k in kats:
        ctr = [int(x, 16) for x in k["ctr"]]
        key = [int(x, 16) for x in k["key"]]
        assert pyoracle.philox(ctr, key) == [int(x, 16) for x in k["out"]]


def _rocrand_bin(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path / "rocrand_philox")
    subprocess.run([hipcc, "-O1", "-std=c++17", "--offload-host-only",
                    os.path.join(HERE, "helpers", "rocrand_philox.cpp"), "-o", out], check=True,
                   capture_output=True)
    return out


def test_philox_matches_rocrand(tmp_path):
    exe = _rocrand_bin(tmp_path)
    g = np.random.default_rng(7)
    rows = g.integers(0, 1 << 32, size=(300, 6), dtype=np.uint64)
    rows[:10, :4] = 0  # small counters as the engine uses them
    text = "\n".join(" ".join(f"{int(x):08x}" for x in r) for r in rows) + "\n"
    res = subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout.split("\n")
    for r, line in zip(rows, res):
        want = [int(x, 16) for x in line.split()]
        assert pyoracle.philox([int(x) for x in r[:4]], [int(x) for x in r[4:]]) == want


def test_numpy_philox_matches_oracle():
    g = np.random.default_rng(11)
    c = g.integers(0, 1 << 32, size=(4, 64), dtype=np.uint64).astype(np.uint32)
    k = g.integers(0, 1 << 32, size=2, dtype=np.uint64)
    got = np.stack(workloads.philox4x32_10(c[0], c[1], c[2], c[3], int(k[0]), int(k[1])), axis=1)
    for i in range(64):
        assert list(got[i]) == pyoracle.philox(c[:, i], k)


def test_stream_word_layout():
    # word j of (stream, gen, slot, tag) is word j&3 of counter block j>>2
    seed = 0x123456789ABCDEF
    for tag in range(5):
        for j in (0, 1, 2, 3, 4, 7, 1023):
            blk = pyoracle.philox([5, 9, 17, (tag << 24) | (j >> 2)], [seed & 0xFFFFFFFF, seed >> 32])
            assert pyoracle.philox_word(seed, 5, 9, 17, tag, j) == blk[j & 3]


def test_key_conversion_is_exact_24_bit():
    assert pyoracle.u32_to_unit(0) == 0.0
    assert pyoracle.u32_to_unit(0xFF) == 0.0
    assert pyoracle.u32_to_unit(0x100) == 2.0 ** -24
    assert pyoracle.u32_to_unit(0xFFFFFFFF) == 1.0 - 2.0 ** -24
    g = np.random.default_rng(3)
    for u in g.integers(0, 1 << 32, size=1000, dtype=np.uint64):
        k = pyoracle.u32_to_unit(int(u))
        assert 0.0 <= k < 1.0
        assert k * 2 ** 24 == (int(u) >> 8)
        assert np.float32(k) == workloads.unit(np.uint32(u))


def test_init_keys_uniformity():
    k = pyoracle.init_keys(42, 0, 64, 1000)
    assert k.min() >= 0.0 and k.max() < 1.0
    assert abs(k.mean() - 0.5) < 0.005
    # 10 equal bins, chi-square with 9 dof well below the 0.1% critical value
    h, _ = np.histogram(k, bins=10, range=(0, 1))
    chi2 = ((h - k.size / 10) ** 2 / (k.size / 10)).sum()
    assert chi2 < 27.9
