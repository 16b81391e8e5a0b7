"""The allocation-free host evaluation (brkga_evaluate_host_noalloc) against
brkga_evaluate_host, through the C ABI: identical fitness and decoded rows
on two calls in a row, the second with fewer rows (the static scratch keeps
the first call's contents), for single- and bi-objective kinds; sizes
beyond the static scratch are rejected with a clear error."""
import ctypes

import numpy as np
import pytest

import brkga_amd
from cases import make_problem
from conftest import random_keys

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 4 << 20


def call(fn, prob, keys, want_decoded=True):
    keys = np.ascontiguousarray(keys, dtype=np.float32)
    rows = keys.shape[0]
    fit = np.full((rows, prob.num_obj), np.nan, np.float32)
    dec = np.full(keys.shape, -7, np.int32) if want_decoded else None
    st = fn(prob.handle, keys.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), rows,
            fit.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
            dec.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if want_decoded else None)
    return st, fit, dec


@pytest.mark.parametrize("kind,n,rows", [("tsp", 300, (400, 37)), ("knap_greedy", 500, (256, 3)),
                                         ("bitsp", 120, (300, 20)), ("tspj", 90, (200, 1))])
def test_noalloc_matches_evaluate_host(engine, kind, n, rows):
    print(f"evaluate_host_noalloc {kind} n={n} rows={rows}", flush=True)
    lib = brkga_amd.lib()
    gp, _ = make_problem(engine, kind, n, inst=n + 5)
    for i, r in enumerate(rows):
        keys = random_keys(1000 + 17 * i + n, r, gp.chrom_len, ties=(i == 1))
        st_ref, f_ref, d_ref = call(lib.brkga_evaluate_host, gp, keys)
        st, f, d = call(lib.brkga_evaluate_host_noalloc, gp, keys)
        assert st_ref == brkga_amd.OK and st == brkga_amd.OK, lib.brkga_last_error()
        assert np.array_equal(f, f_ref), f"fitness differs ({kind}, call {i}, {r} rows)"
        assert np.array_equal(d, d_ref), f"decoded rows differ ({kind}, call {i}, {r} rows)"
        # fitness only: same values, decoded untouched
        st, f2, _ = call(lib.brkga_evaluate_host_noalloc, gp, keys, want_decoded=False)
        assert st == brkga_amd.OK and np.array_equal(f2, f_ref)


def test_noalloc_rejects_rows_beyond_scratch(engine):
    lib = brkga_amd.lib()
    gp, _ = make_problem(engine, "tsp", 1000, inst=3)
    rows = SCRATCH_BYTES // (8 * 1000) + 1
    keys = random_keys(5, rows, 1000)
    st, _, _ = call(lib.brkga_evaluate_host_noalloc, gp, keys)
    assert st == brkga_amd.ERR_INVALID_ARG
    assert b"static scratch" in lib.brkga_last_error()
    # the largest call that fits still works, and agrees with the allocating form
    fit_rows = (SCRATCH_BYTES - 256) // (8 * 1000) - 1
    keys = random_keys(6, fit_rows, 1000)
    st, f, d = call(lib.brkga_evaluate_host_noalloc, gp, keys)
    st_ref, f_ref, d_ref = call(lib.brkga_evaluate_host, gp, keys)
    assert st == brkga_amd.OK and st_ref == brkga_amd.OK
    assert np.array_equal(f, f_ref) and np.array_equal(d, d_ref)
