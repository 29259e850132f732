"""GPU tests of the request loop (hg_server_*): every answer equals the
single-graph engine's on that graph alone, whatever the mix, batching,
lane count or memory budget; admission respects the budget."""
import dataclasses
import sys
import threading
import time

import numpy as np
import pytest

from pargridcpu_amd.graph import MultilayerGraph, batch_spec
from pargridcpu_amd.hipgrid import REQ_DECOMPOSE, REQ_DENSEST, Engine, HgError, Server
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


def check(srv, rid, mg, kind, beta):
    info = srv.wait(rid)
    assert info.status == 0
    ci, kmax = srv.result(rid)
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
