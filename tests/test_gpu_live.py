"""GPU tests of the device work queue shared by batches (hg_live): batches
submitted while others run are taken by the workgroups already on the GPU;
every graph's core index equals the oracle's CoreIndex restatement."""
import sys
import threading

import numpy as np
import pytest

from oracle import oracle as O
from pargridcpu_amd.graph import MultilayerGraph, batch_spec
from pargridcpu_amd.hipgrid import Batch, Live, REQ_DECOMPOSE, REQ_DENSEST, Server
from tests import graphs as G

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def graphs(count, base, n_lo=200, n_hi=3000):
    return [MultilayerGraph.generate(batch_spec(g, base=base, n_lo=n_lo, n_hi=n_hi)) for g in range(count)]


def check_batch(b, gs):
    for i, mg in enumerate(gs):
        ref = O.coreindex(mg.n, mg.layers())
        assert np.array_equal(b.core_index(i), ref), f"graph {i}"
        cl, kmax = b.cells(i)
        assert np.array_equal(kmax, ref.max(axis=1)), f"graph {i}"


def test_live_single_batch_and_resubmit():
    gs = graphs(12, 0xE1) + [G.graph(10, [[], []]), G.graph(1, [[]])]
    say("case live single batch")
    with Live() as q, Batch(gs) as b:
        for rep in range(3):
            b.submit_live(q)
            b.wait_live()
            check_batch(b, gs)
        st = q.stats()
        assert st["items_done"] == 3 * b.info()["items"]


def test_live_admission_into_running_launch():
    """A first batch with more items than the GPU holds workgroups keeps every
    CU busy; batches submitted meanwhile are taken by its workgroups (their
    own launches cannot start until CUs free up), and every result is exact."""
    big = graphs(160, 0xE2, n_lo=1000, n_hi=8000)
    small = [graphs(4, 0xE3 + i) for i in range(3)]
    say(f"case live admission: {len(big)} + 3x4 graphs")
    with Live() as q:
        b0 = Batch(big)
        bs = [Batch(g) for g in small]
        b0.submit_live(q)
        for b in bs:
            b.submit_live(q)
        b0.wait_live()
        taken = [b.wait_live() for b in bs]
        check_batch(b0, big)
        for b, g in zip(bs, small):
            check_batch(b, g)
        assert sum(taken) > 0, "no later item was taken by an earlier launch"
        st = q.stats()
        assert st["items_taken_running"] >= sum(taken)
        for b in [b0] + bs:
            b.close()


def test_live_concurrent_submitters():
    """Several host threads submit and wait on one queue at once."""
    sets = [graphs(6, 0xE8 + t, n_hi=5000) for t in range(4)]
    say("case live concurrent submitters")
    errors = []
    with Live(block=512, unroll=4, blocks_per_cu=2) as q:
        def worker(gs):
            try:
                with Batch(gs) as b:
                    for _ in range(3):
                        b.submit_live(q)
                        b.wait_live()
                        check_batch(b, gs)
            except Exception as e:  # reported below
                errors.append(e)

        ts = [threading.Thread(target=worker, args=(gs,)) for gs in sets]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
    assert not errors, errors


def test_server_device_queue_mode():
    """Serving with device_queue: small requests go straight to the shared
    queue; every answer (decomposition and densest) equals the oracle's."""
    gs = graphs(24, 0xEE, n_hi=20000) + [G.graph(10, [[], []])]
    say("case server device queue")
    with Server(lanes=4, max_batch=4, device_queue=True, large_entries=1 << 40) as srv:
        ids = [srv.submit(g, REQ_DENSEST if i % 5 == 0 else REQ_DECOMPOSE, 1.0) for i, g in enumerate(gs)]
        for i, (rid, g) in enumerate(zip(ids, gs)):
            info = srv.wait(rid)
            ci, kmax = srv.result(rid)
            ref = O.coreindex(g.n, g.layers())
            assert np.array_equal(ci, ref), f"request {i}"
            if i % 5 == 0:
                w = O.wds(g.n, g.layers(), 1.0)
                assert np.float32(info.wds.max_density) == w["max_density"]
                assert (info.wds.res_k, info.wds.res_lmd, info.wds.res_len) == (w["res_k"], w["res_lmd"], w["res_len"])
        st = srv.stats()
        assert st["completed"] == len(gs) and st["failed"] == 0
