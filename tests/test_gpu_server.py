"""GPU tests of the request loop (hg_server_*): every answer equals the
oracle's (CoreIndex restatement for the core index, WdsList restatement for
densest requests) and the single-graph engine's on that graph alone,
whatever the mix, batching, lane count or memory budget; admission respects
the budget."""
import dataclasses
import sys
import threading
import time

import numpy as np
import pytest

from oracle import oracle as O
from pargridcpu_amd.graph import MultilayerGraph, batch_spec, large_request_spec
from pargridcpu_amd.hipgrid import REQ_DECOMPOSE, REQ_DENSEST, Engine, HgError, Server, lib
from tests import graphs as G

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def request_graphs(count, base):
    gs = [MultilayerGraph.generate(batch_spec(g, base=base, n_lo=300, n_hi=20000)) for g in range(count)]
    gs.append(G.graph(10, [[], []]))  # no edges
    return gs


def alone(mg, kind, beta):
    with Engine.from_graph(mg) as eng:
        eng.decompose()
        ci = eng.core_index()
        _, kmax = eng.cells()
        w = eng.densest(beta) if kind == REQ_DENSEST else None
    return ci, kmax, w


_ORACLE = {}


def oracle_answer(mg, kind, beta):
    """Core index from the oracle's CoreIndex restatement (one lmd per host
    thread) and, for densest requests, the oracle's WdsList result; cached
    per graph."""
    key = (id(mg), kind, beta)
    if key not in _ORACLE:
        ci = O.coreindex_parallel(mg.n, mg.layers())
        w = O.wds(mg.n, mg.layers(), beta) if kind == REQ_DENSEST else None
        _ORACLE[key] = (mg, ci, w)  # mg kept so id() stays unique
    return _ORACLE[key][1:]


def check(srv, rid, mg, kind, beta, with_alone=True):
    info = srv.wait(rid)
    assert info.status == 0
    ci, kmax = srv.result(rid)
    oci, ow = oracle_answer(mg, kind, beta)
    assert np.array_equal(ci, oci), f"request {rid}: core index differs from the oracle"
    assert np.array_equal(kmax, oci.max(axis=1))
    if kind == REQ_DENSEST:
        w = info.wds
        assert np.float32(w.max_density) == ow["max_density"], f"request {rid}"
        assert (w.res_k, w.res_lmd, w.res_len) == (ow["res_k"], ow["res_lmd"], ow["res_len"])
    if with_alone:
        eci, ekmax, ew = alone(mg, kind, beta)
        assert np.array_equal(ci, eci), f"request {rid}"
        assert np.array_equal(kmax, ekmax)
        if kind == REQ_DENSEST:
            w = info.wds
            assert np.float32(w.max_density) == np.float32(ew.max_density)
            assert (w.res_k, w.res_lmd, w.res_len) == (ew.res_k, ew.res_lmd, ew.res_len)
    return info


@pytest.mark.parametrize("lanes,max_batch,plain,min_batch", [(3, 8, False, 1), (1, 64, False, 0), (1, 1, True, 0),
                                                               (4, 16, False, 4)])
def test_mix_equals_alone(lanes, max_batch, plain, min_batch):
    gs = request_graphs(20, 0xD1)
    kinds = [REQ_DENSEST if i % 5 == 0 else REQ_DECOMPOSE for i in range(len(gs))]
    betas = [1.0 if i % 2 else 0.5 for i in range(len(gs))]
    say(f"case server lanes={lanes} max_batch={max_batch} plain={plain} min_batch={min_batch}")
    # large_entries low enough that the two largest graphs run alone on the
    # single-graph engine while the rest are batched (min_batch 1: whenever
    # any small request waits) or run alone when few wait
    big = sorted(g.directed_entries() for g in gs)[-3]
    with Server(lanes=lanes, max_batch=max_batch, plain=plain, large_entries=big, min_batch=min_batch) as srv:
        ids = []
        for i, g in enumerate(gs):  # staggered arrivals: later ones join batches formed meanwhile
            ids.append(srv.submit(g, kinds[i], betas[i]))
            if i % 4 == 3:
                time.sleep(0.002)
        sizes = []
        for rid, g, k, b in zip(ids, gs, kinds, betas):
            sizes.append(check(srv, rid, g, k, b).batch_size)
        st = srv.stats()
        assert st["completed"] == len(gs) and st["failed"] == 0
        assert st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
        if plain:
            assert set(sizes) == {1}
        for rid in ids:
            srv.release(rid)


def test_budget_admission():
    gs = request_graphs(6, 0xD2)
    say("case server budget")
    # a budget that admits a few small requests at a time, and one request
    # (about 2 GB of device state) that can never fit
    budget = 512 << 20
    with Server(lanes=2, max_batch=8, hbm_budget=budget) as srv:
        huge = MultilayerGraph.generate(dataclasses.replace(batch_spec(0, base=0xD3), n=2_000_000, L=8,
                                                            m_lo=10_000_000, m_hi=10_000_000))
        ids = [srv.submit(g) for g in gs]
        bad = srv.submit(huge)
        for rid, g in zip(ids, gs):
            check(srv, rid, g, REQ_DECOMPOSE, 1.0)
        with pytest.raises(HgError) as e:
            srv.wait(bad)
        assert e.value.code == -3
        st = srv.stats()
        assert 0 < st["peak_bytes"] <= budget
        assert st["failed"] == 1


def test_concurrent_submitters():
    gs = request_graphs(12, 0xD4)
    say("case server concurrent submitters")
    with Server(lanes=4, max_batch=4) as srv:
        ids = [None] * len(gs)

        def worker(lo):
            for i in range(lo, len(gs), 3):
                ids[i] = srv.submit(gs[i])

        ts = [threading.Thread(target=worker, args=(k,)) for k in range(3)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        srv.drain()
        for rid, g in zip(ids, gs):
            check(srv, rid, g, REQ_DECOMPOSE, 1.0)


def test_concurrent_densest_batches():
    """Densest requests in several lanes at once: each lane uploads its
    batch's item list while the others allocate and free host memory."""
    gs = request_graphs(24, 0xD5)
    say("case server concurrent densest batches")
    with Server(lanes=4, max_batch=3) as srv:
        ids = [srv.submit(g, REQ_DENSEST, 1.0) for g in gs]
        for rid, g in zip(ids, gs):
            check(srv, rid, g, REQ_DENSEST, 1.0)


def test_large_requests_in_the_mix():
    """The serving workload's large request (n = 5M, L = 12, 60M sampled
    edges) twice, submitted among small decomposition and densest requests
    with large_entries below its size: the large ones run alone on the
    single-graph engine while the small ones are batched around them; every
    answer equals the oracle's."""
    large = MultilayerGraph.generate(large_request_spec(0))
    small = request_graphs(16, 0xD6)
    say(f"case server large requests n={large.n} L={large.L} D={large.directed_entries()} among 17 small")
    say("  oracle for the large graph (one lmd per host thread)")
    oracle_answer(large, REQ_DECOMPOSE, 1.0)
    with Server(lanes=3, max_batch=8, min_batch=2, large_entries=2_000_000) as srv:
        plan = [(g, REQ_DENSEST if i % 4 == 1 else REQ_DECOMPOSE, 1.0) for i, g in enumerate(small)]
        plan.insert(3, (large, REQ_DECOMPOSE, 1.0))
        plan.insert(11, (large, REQ_DECOMPOSE, 1.0))
        ids = [srv.submit(g, k, b) for g, k, b in plan]
        sizes = {}
        for rid, (g, k, b) in zip(ids, plan):
            info = check(srv, rid, g, k, b, with_alone=g is not large)
            sizes.setdefault(g is large, []).append(info.batch_size)
        assert sizes[True] == [1, 1]
        st = srv.stats()
        assert st["completed"] == len(plan) and st["failed"] == 0


@pytest.mark.parametrize("reuse", [True, False])
def test_resident_graphs(reuse):
    """Registered graphs stay resident: repeat decompositions and densest
    searches on them (by handle, nothing uploaded) equal the oracle, mixed
    with copied-in requests, whether the resident decomposition is reused
    or recomputed; their bytes count against the budget; a graph with
    queued requests cannot be unregistered."""
    gs = request_graphs(6, 0xD7)
    extra = request_graphs(4, 0xD8)
    say(f"case server resident graphs reuse={reuse}")
    with Server(lanes=3, max_batch=4, min_batch=2, reuse_results=reuse) as srv:
        gids = [srv.register(g) for g in gs]
        st = srv.stats()
        assert st["resident_graphs"] == len(gs) and st["resident_bytes"] > 0
        plan = []
        for rep in range(3):
            for gid, g in zip(gids, gs):
                kind = REQ_DENSEST if (rep + gid) % 3 == 0 else REQ_DECOMPOSE
                plan.append((srv.submit_resident(gid, kind, 1.0 if rep else 0.5), g, kind, 1.0 if rep else 0.5))
        for g in extra:
            plan.append((srv.submit(g, REQ_DECOMPOSE, 1.0), g, REQ_DECOMPOSE, 1.0))
        for rid, g, kind, beta in plan:
            check(srv, rid, g, kind, beta, with_alone=False)
        srv.drain()
        for gid in gids:
            srv.unregister(gid)
        st = srv.stats()
        assert st["completed"] == len(plan) and st["failed"] == 0
        assert st["resident_graphs"] == 0 and st["resident_bytes"] == 0
        with pytest.raises(HgError) as e:
            srv.unregister(gids[0])
        assert e.value.code == -1


def test_register_respects_budget():
    gs = request_graphs(2, 0xD9)
    say("case server register over budget")
    with Server(lanes=1, hbm_budget=64 << 20) as srv:
        big = MultilayerGraph.generate(dataclasses.replace(batch_spec(0, base=0xDA), n=1_000_000, L=6,
                                                           m_lo=4_000_000, m_hi=4_000_000))
        with pytest.raises(HgError) as e:
            srv.register(big)
        assert e.value.code == -3
        gid = srv.register(gs[0])
        check(srv, srv.submit_resident(gid), gs[0], REQ_DECOMPOSE, 1.0)


def test_cancel_queued_requests():
    """With one lane busy on a large request, queued requests can be
    cancelled: they end with HG_ERR_CANCELLED and the rest are answered."""
    big = MultilayerGraph.generate(dataclasses.replace(batch_spec(0, base=0xDB), n=2_000_000, L=8,
                                                       m_lo=6_000_000, m_hi=6_000_000))
    gs = request_graphs(5, 0xDC)
    say("case server cancel")
    with Server(lanes=1, plain=True) as srv:
        gid = srv.register(gs[0])
        first = srv.submit(big)
        ids = [srv.submit(g) for g in gs]
        on_resident = srv.submit_resident(gid)
        with pytest.raises(HgError) as e:  # queued behind the large request
            srv.unregister(gid)
        assert e.value.code == -5
        srv.cancel(on_resident)
        srv.unregister(gid)
        cancelled = []
        for rid in ids[1::2]:
            try:
                srv.cancel(rid)
                cancelled.append(rid)
            except HgError as e:  # it had started already (the lane was quicker)
                assert e.code == -5
        assert cancelled, "the lane finished a 2M-vertex request before the next submit"
        for rid, g in zip(ids, gs):
            if rid in cancelled:
                with pytest.raises(HgError) as e:
                    srv.wait(rid)
                assert e.value.code == -7
                srv.cancel(rid)  # again: a no-op
            else:
                check(srv, rid, g, REQ_DECOMPOSE, 1.0, with_alone=False)
        assert srv.wait(first).status == 0
        srv.drain()
        st = srv.stats()
        assert st["cancelled"] == len(cancelled) + 1
        assert st["completed"] == len(gs) + 1 - len(cancelled)
        with pytest.raises(HgError) as e:
            srv.cancel(first)
        assert e.value.code == -5


def test_injected_allocation_failure():
    """A device allocation that fails (injected) fails only the request or
    batch that made it, with HG_ERR_NOMEM; the loop keeps serving and the
    other answers equal the oracle."""
    gs = request_graphs(8, 0xDD)
    say("case server injected allocation failure")
    L = lib()
    try:
        with Server(lanes=1, plain=True) as srv:
            assert L.hg_debug_fail_alloc(1) == 0  # the next allocation: the first request's set-up
            ids = [srv.submit(g) for g in gs[:3]]
            with pytest.raises(HgError) as e:
                srv.wait(ids[0])
            assert e.value.code == -3
            for rid, g in zip(ids[1:], gs[1:3]):
                check(srv, rid, g, REQ_DECOMPOSE, 1.0, with_alone=False)
        with Server(lanes=2, max_batch=8, min_batch=2, large_entries=1 << 40) as srv:
            assert L.hg_debug_fail_alloc(6) == 0  # somewhere inside the first batch's set-up
            ids = [srv.submit(g) for g in gs]
            failed = 0
            for rid, g in zip(ids, gs):
                try:
                    check(srv, rid, g, REQ_DECOMPOSE, 1.0, with_alone=False)
                except HgError as e:
                    assert e.code == -3
                    failed += 1
            assert failed >= 1
            later = srv.submit(gs[0])
            check(srv, later, gs[0], REQ_DECOMPOSE, 1.0, with_alone=False)
    finally:
        L.hg_debug_fail_alloc(0)


def test_resident_graphs_on_the_device_queue():
    """With device_queue, small registered graphs are kept as one-graph
    batches whose items join the shared device queue; large ones on the
    single-graph engine; every answer equals the oracle's."""
    gs = request_graphs(8, 0xDE)
    say("case server resident graphs on the device queue")
    big = sorted(g.directed_entries() for g in gs)[-2]
    with Server(lanes=4, max_batch=4, device_queue=True, large_entries=big, reuse_results=False) as srv:
        gids = [srv.register(g) for g in gs]
        plan = []
        for rep in range(2):
            for gid, g in zip(gids, gs):
                kind = REQ_DENSEST if (rep + gid) % 3 == 0 else REQ_DECOMPOSE
                plan.append((srv.submit_resident(gid, kind, 1.0), g, kind, 1.0))
        for rid, g, kind, beta in plan:
            check(srv, rid, g, kind, beta, with_alone=False)
        srv.drain()
        for gid in gids:
            srv.unregister(gid)
