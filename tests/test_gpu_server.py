"""Serving form: requests admitted by HBM budget, batched while resident,
retired at their generation deadline.

The bar: every request's best fitness and decoded solution are identical to
those of the same population evolved alone (brkga_evolve) for the same
number of generations, whatever else shared the device with it.
"""
import time

import numpy as np
import pytest

import brkga_amd
from cases import make_problem

pytestmark = pytest.mark.gpu

REQUESTS = [
    # kind, n, P, E, M, rho, islands, interval, migrants, generations
    ("tsp", 25, 64, 12, 6, 0.7, 1, 0, 0, 9),
    ("tsp", 300, 128, 20, 20, 0.7, 2, 3, 2, 6),
    ("knap_greedy", 150, 64, 12, 6, 0.7, 1, 0, 0, 11),
    ("bitsp", 40, 64, 0, 0, 0.7, 1, 0, 0, 5),
    ("tsp", 1200, 64, 10, 10, 0.7, 1, 0, 0, 4),
    ("tsp_matrix", 50, 48, 9, 5, 0.65, 1, 0, 0, 7),
    ("knap_threshold", 90, 64, 12, 6, 0.7, 1, 0, 0, 0),
    ("tspj", 40, 64, 12, 6, 0.75, 1, 0, 0, 1),
    ("tspj_ls", 35, 64, 12, 6, 0.75, 2, 2, 1, 7),
    ("tsp", 64, 64, 10, 10, 0.7, 1, 0, 0, 13),
    ("bitsp", 120, 128, 0, 0, 0.7, 1, 0, 0, 3),
    ("knap_greedy", 700, 128, 20, 12, 0.7, 1, 0, 0, 5),
    ("tsp", 5000, 32, 5, 5, 0.7, 1, 0, 0, 2),
    ("jobshop", 120, 64, 12, 6, 0.7, 1, 0, 0, 6),
    ("cont_sphere", 40, 64, 12, 6, 0.7, 1, 0, 0, 5),
    ("vrp_days", 60, 64, 12, 6, 0.7, 1, 0, 0, 4),
    ("zdt1", 30, 32, 0, 0, 0.5, 1, 0, 0, 7),
    ("movrp", 31, 32, 0, 0, 0.5, 1, 0, 0, 3),
]


def solo_best(gp, prm, G):
    pop = brkga_amd.Population(gp, prm)
    pop.initialize()
    pop.evolve(G)
    fit, _, dec = pop.best()
    pop.close()
    return fit, dec


def build_requests(engine):
    out = []
    for i, (kind, n, P, E, M, rho, I, iv, k, G) in enumerate(REQUESTS):
        ls = (2, 4, 5) if kind == "tspj_ls" else (0, 0, 0)
        gp, _ = make_problem(engine, "tspj" if kind == "tspj_ls" else kind, n, inst=200 + i)
        prm = brkga_amd.Params(P, E, M, rho, I, iv, k, 0x5E4E + 11 * i, *ls)
        out.append((gp, prm, G))
    return out


def check(result, ref, ctx):
    fit, dec, info = result
    rfit, rdec = ref
    assert info["status"] == 0, ctx
    assert np.array_equal(np.asarray(fit), np.asarray(rfit)), f"best fitness differs ({ctx})"
    assert np.array_equal(dec, rdec), f"decoded solution differs ({ctx})"


@pytest.mark.parametrize("budget_mib,max_resident,ticks,tune", [(512, 0, 3, 0), (24, 3, 1, 0), (512, 2, 5, 1)])
def test_server_results_match_solo(engine, budget_mib, max_resident, ticks, tune):
    print(f"server mix budget={budget_mib} MiB max_resident={max_resident} ticks={ticks} tune={tune}", flush=True)
    reqs = build_requests(engine)
    refs = [solo_best(gp, prm, G) for gp, prm, G in reqs]
    srv = brkga_amd.Server(engine, hbm_budget_bytes=budget_mib << 20, max_resident=max_resident,
                           ticks_per_round=ticks, autotune_reps=tune)
    ids = []
    for i, (gp, prm, G) in enumerate(reqs):
        ids.append(srv.submit(gp, prm, G))
        if i % 4 == 3:
            time.sleep(0.002)  # later requests join while earlier ones run
    results = [srv.wait(r, timeout_ms=60000) for r in ids]
    st = srv.stats()
    srv.stop()
    for i, (res, ref) in enumerate(zip(results, refs)):
        check(res, ref, f"request {i} {REQUESTS[i][:2]}")
        assert res[2]["generations"] == REQUESTS[i][-1]
        assert res[2]["submit_us"] <= res[2]["admit_us"] <= res[2]["done_us"]
    assert st["completed"] == len(reqs) and st["submitted"] == len(reqs)
    assert st["peak_bytes"] <= st["budget_bytes"]
    assert 0 <= st["lanes"] <= 4  # 0: every round ran the requests on their own streams


def test_server_alone_equals_in_mix(engine):
    """One request served alone and served among others: same result."""
    print("server alone vs mix", flush=True)
    reqs = build_requests(engine)
    gp, prm, G = reqs[1]
    srv = brkga_amd.Server(engine, hbm_budget_bytes=256 << 20)
    alone = srv.wait(srv.submit(gp, prm, G), timeout_ms=60000)
    ids = [srv.submit(q, p, g) for q, p, g in reqs]
    mixed = [srv.wait(r, timeout_ms=60000) for r in ids]
    srv.stop()
    check(mixed[1], (alone[0], alone[1]), "request 1 alone vs in the mix")


def test_server_long_run_with_joins_matches_solo(engine):
    """A request alone for many rounds (the single-population launches of a
    lone request), then two more joining while it runs (batched ticks), then
    alone again: every result still equals its solo run."""
    print("server long run with joins", flush=True)
    reqs = build_requests(engine)
    picks = [(reqs[1][0], reqs[1][1], 3000), (reqs[16][0], reqs[16][1], 25), (reqs[0][0], reqs[0][1], 30)]
    refs = [solo_best(gp, prm, G) for gp, prm, G in picks]
    srv = brkga_amd.Server(engine, hbm_budget_bytes=512 << 20, ticks_per_round=2)
    first = srv.submit(*picks[0])
    time.sleep(0.005)  # rounds with one member, then two requests join
    ids = [first] + [srv.submit(*p) for p in picks[1:]]
    results = [srv.wait(r, timeout_ms=60000) for r in ids]
    st = srv.stats()
    srv.stop()
    for i, (res, ref) in enumerate(zip(results, refs)):
        check(res, ref, f"long run with joins, request {i}")
        assert res[2]["generations"] == picks[i][2]
    assert st["completed"] == 3
    # the two short requests ran while the long one was resident
    assert all(res[2]["done_us"] < results[0][2]["done_us"] for res in results[1:])


def test_server_lone_requests_in_sequence(engine):
    """Requests that each find the device empty (single-population rounds)
    give their solo results, one after the other, for every kind."""
    print("server lone requests", flush=True)
    reqs = build_requests(engine)
    srv = brkga_amd.Server(engine, hbm_budget_bytes=512 << 20)
    for i in (0, 2, 3, 6, 8, 13, 16):
        gp, prm, G = reqs[i]
        res = srv.wait(srv.submit(gp, prm, G), timeout_ms=60000)
        check(res, solo_best(gp, prm, G), f"lone request {i} {REQUESTS[i][:2]}")
        assert res[2]["generations"] == G
    srv.stop()


def test_server_rejects_and_reports(engine):
    print("server rejection / wait semantics", flush=True)
    gp, _ = make_problem(engine, "tsp", 2000, inst=7)
    prm = brkga_amd.Params(1024, 150, 150, 0.7, 1, 0, 0, 3)
    srv = brkga_amd.Server(engine, hbm_budget_bytes=1 << 20)
    with pytest.raises(brkga_amd.BrkgaError) as e:
        srv.submit(gp, prm, 10)
    assert e.value.status == brkga_amd.ERR_OOM
    with pytest.raises(brkga_amd.BrkgaError) as e:
        srv.wait(12345, timeout_ms=0)
    assert e.value.status == brkga_amd.ERR_NOT_FOUND
    small, _ = make_problem(engine, "tsp", 30, inst=8)
    rid = srv.submit(small, brkga_amd.Params(32, 6, 4, 0.7, 1, 0, 0, 4), 3)
    fit, dec, info = srv.wait(rid, timeout_ms=60000)
    assert sorted(dec.tolist()) == list(range(30))
    with pytest.raises(brkga_amd.BrkgaError) as e:
        srv.wait(rid, timeout_ms=0)
    assert e.value.status == brkga_amd.ERR_NOT_FOUND
    st = srv.stats()
    assert st["rejected"] == 1 and st["completed"] == 1
    srv.stop()


def test_server_cancel(engine):
    """A cancelled pending request and a cancelled resident request end with
    ERR_CANCELLED; every other request still equals its solo run; the HBM of
    the cancelled ones returns to the budget."""
    print("server cancel: one resident at a time, long first request", flush=True)
    reqs = build_requests(engine)[:6]
    long_gp, _ = make_problem(engine, "tsp", 400, inst=777)
    long_prm = brkga_amd.Params(256, 38, 38, 0.7, 1, 0, 0, 0xC0FFEE)
    refs = [solo_best(gp, prm, G) for gp, prm, G in reqs]
    srv = brkga_amd.Server(engine, hbm_budget_bytes=512 << 20, max_resident=1, ticks_per_round=2)
    long_id = srv.submit(long_gp, long_prm, 1_000_000)  # runs until cancelled
    ids = [srv.submit(gp, prm, G) for gp, prm, G in reqs]
    srv.cancel(ids[2])  # still pending: max_resident = 1 and the long request holds the device
    time.sleep(0.05)
    assert srv.stats()["resident"] == 1
    srv.cancel(long_id)  # resident: leaves at the next round
    for k, (rid, ref) in enumerate(zip(ids, refs)):
        if k == 2:
            with pytest.raises(brkga_amd.BrkgaError) as e:
                srv.wait(rid, timeout_ms=60000)
            assert e.value.status == brkga_amd.ERR_CANCELLED
            continue
        check(srv.wait(rid, timeout_ms=60000), ref, f"request {k} after cancels")
    with pytest.raises(brkga_amd.BrkgaError) as e:
        srv.wait(long_id, timeout_ms=60000)
    assert e.value.status == brkga_amd.ERR_CANCELLED
    st = srv.stats()
    assert st["bytes_in_use"] == 0 and st["resident"] == 0 and st["pending"] == 0
    # unknown ids; a collected id is unknown too
    with pytest.raises(brkga_amd.BrkgaError) as e:
        srv.cancel(10**9)
    assert e.value.status == brkga_amd.ERR_NOT_FOUND
    with pytest.raises(brkga_amd.BrkgaError) as e:
        srv.cancel(ids[0])
    assert e.value.status == brkga_amd.ERR_NOT_FOUND
    srv.stop()
