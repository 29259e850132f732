"""GPU tests of batched decomposition (hg_batch_*): every graph of a batch
gets exactly the results of the oracle and of the single-graph engine on
that graph alone, for every launch shape, on every replay of the captured
launch."""
import ctypes as C
import dataclasses
import sys
import threading

import numpy as np
import pytest

from oracle import oracle as O
from pargridcpu_amd.graph import MultilayerGraph, batch_spec
from pargridcpu_amd.hipgrid import Batch, Engine, HgError, lib
from tests import graphs as G

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def mixed_graphs(count=24, n_hi=6000, base=0xB7):
    """Seeded graphs of the batched workload's shape distribution at small n,
    plus hand-built edge cases."""
    gs = [MultilayerGraph.generate(batch_spec(g, base=base, n_lo=200, n_hi=n_hi)) for g in range(count)]
    n = 40
    es = G.clique(range(6)) + G.star(6, range(7, 15)) + G.cycle(range(15, 25)) + G.path(range(25, 33))
    gs.append(G.graph(n, [es, es, es]))                        # identical layers
    gs.append(G.graph(10, [[], []]))                           # no edges at all
    gs.append(G.graph(1, [[]]))                                # one vertex
    gs.append(G.graph(5001, [G.star(0, range(1, 5001))] * 2))  # one hub
    gs.append(G.graph(10, [G.clique(range(5)), G.clique(range(5, 10))]))  # disjoint layers
    # more layers than the batched level-0 pull takes: the item pulls its own rows
    gs.append(MultilayerGraph.generate(dataclasses.replace(batch_spec(count, base=base, n_lo=200, n_hi=n_hi), L=20)))
    return gs


def engine_results(mg, beta=None):
    with Engine.from_graph(mg) as eng:
        eng.decompose()
        ci = eng.core_index()
        cl, kmax = eng.cells()
        w = eng.densest(beta) if beta is not None else None
    return ci, cl, kmax, w


def test_batch_matches_oracle_and_engine():
    gs = mixed_graphs()
    say(f"case batch mixed: {len(gs)} graphs")
    with Batch(gs) as b:
        info = b.info()
        assert info["graphs"] == len(gs)
        b.decompose()
        for i, mg in enumerate(gs):
            ci = b.core_index(i)
            ref = O.coreindex(mg.n, mg.layers())
            assert np.array_equal(ci, ref), f"graph {i}"
            cl, kmax = b.cells(i)
            assert np.array_equal(kmax, ref.max(axis=1) if mg.n else kmax)
            eci, ecl, ekmax, _ = engine_results(mg)
            assert np.array_equal(ci, eci)
            assert np.array_equal(kmax, ekmax)
            w = min(cl.shape[1], ecl.shape[1])
            assert np.array_equal(cl[:, :w], ecl[:, :w])


def test_batch_few_layers():
    """Every graph with at most 8 layers (the narrow level-0 counters), hubs
    of several hundred entries per layer included."""
    gs = [MultilayerGraph.generate(dataclasses.replace(batch_spec(g, base=0xBA, n_lo=300, n_hi=5000), L=2 + g % 7))
          for g in range(10)]
    gs.append(G.graph(3001, [G.star(0, range(1, 3001)), G.star(0, range(1, 1500)), G.star(1, range(2, 700))]))
    # 20 layers of which 5 hold edges: few items, many rows per vertex
    rng = np.random.default_rng(0xBB)
    gs.append(G.graph(2000, [[tuple(e) for e in rng.integers(0, 2000, size=(6000, 2))] if l % 4 == 0 else []
                             for l in range(20)]))
    # a layer degree of 2^16 or more: one degree per LDS hub word
    big = 70_001
    gs.append(G.graph(big, [G.star(0, range(1, big)) + [tuple(e) for e in rng.integers(0, big, size=(50_000, 2))],
                            G.cycle(range(big)) + G.clique(range(40))]))
    say(f"case batch few layers: {len(gs)} graphs")
    with Batch(gs) as b:
        b.decompose()
        for i, mg in enumerate(gs):
            assert np.array_equal(b.core_index(i), O.coreindex(mg.n, mg.layers())), f"graph {i}"


def test_stale_item_state(monkeypatch):
    """Item state that starts as all-ones words instead of zeros (HG_POISON=1):
    rows of vertices that take no part in an item are never written, and
    with an odd number of layers a 16-bit LDS degree word pairs such a row
    with a valid one; results stay exact."""
    monkeypatch.setenv("HG_POISON", "1")
    gs = [MultilayerGraph.generate(dataclasses.replace(batch_spec(g, base=0xBE, n_lo=300, n_hi=6000), L=L))
          for g, L in enumerate([3, 5, 7, 9, 11, 13, 15, 4])]
    say(f"case batch stale item state: {len(gs)} graphs")
    for block in (1024, 512):
        with Batch(gs) as b:
            b.set_params(block=block, unroll=4, blocks_per_cu=1, order=0, use_graph=1)
            b.decompose()
            for i, mg in enumerate(gs):
                assert np.array_equal(b.core_index(i), O.coreindex(mg.n, mg.layers())), f"block {block} graph {i}"


@pytest.mark.parametrize("block,unroll,bpc,order,use_graph", [(256, 2, 1, 0, 1), (512, 4, 2, 1, 1), (1024, 4, 1, 0, 0),
                                                               (256, 4, 8, 1, 0)])
def test_launch_shapes_agree(block, unroll, bpc, order, use_graph):
    gs = mixed_graphs(count=12, base=0xB8)
    refs = [O.coreindex(mg.n, mg.layers()) for mg in gs]
    say(f"case batch shape block={block} unroll={unroll} bpc={bpc} order={order} graph={use_graph}")
    with Batch(gs) as b:
        b.set_params(block=block, unroll=unroll, blocks_per_cu=bpc, order=order, use_graph=use_graph)
        for rep in range(3):  # replays of the captured launch, with host work between them
            b.decompose()
            for i in range(len(gs)):
                assert np.array_equal(b.core_index(i), refs[i]), f"replay {rep} graph {i}"


def test_large_graph_global_bitmap():
    """n above the LDS bitmap capacity (262144 vertices): the item keeps its
    liveness bits in global memory."""
    spec = dataclasses.replace(batch_spec(3, base=0xB9), n=300_000, m_lo=600_000, m_hi=600_000, L=3)
    gs = [MultilayerGraph.generate(spec), MultilayerGraph.generate(batch_spec(1, base=0xB9, n_hi=3000))]
    say("case batch large graph n=300000")
    with Batch(gs) as b:
        b.decompose()
        for i, mg in enumerate(gs):
            assert np.array_equal(b.core_index(i), O.coreindex(mg.n, mg.layers()))


@pytest.mark.parametrize("beta", [1.0, 0.5, 2.0])
def test_batch_densest(beta):
    gs = mixed_graphs(count=10, base=0xBA)
    say(f"case batch densest beta={beta}")
    with Batch(gs) as b:
        b.decompose()
        sel = list(range(len(gs)))
        ws = b.densest(sel, beta)
        for i in sel:
            _, _, _, w = engine_results(gs[i], beta)
            got = ws[i]
            assert np.float32(got.max_density) == np.float32(w.max_density), f"graph {i}"
            assert (got.res_k, got.res_lmd, got.res_len) == (w.res_k, w.res_lmd, w.res_len), f"graph {i}"
            ref = O.wds(gs[i].n, gs[i].layers(), beta)
            assert np.float32(got.max_density) == ref["max_density"]


def test_tune_keeps_results():
    gs = mixed_graphs(count=8, base=0xBB)
    refs = [O.coreindex(mg.n, mg.layers()) for mg in gs]
    say("case batch tune")
    with Batch(gs) as b:
        best, ms = b.tune(reps=1)
        assert best["block"] in (256, 512, 1024) and ms > 0
        b.decompose()
        for i in range(len(gs)):
            assert np.array_equal(b.core_index(i), refs[i])


def test_batch_errors():
    gs = mixed_graphs(count=2, base=0xBC)
    with Batch(gs) as b:
        with pytest.raises(HgError) as e:
            b.core_index(0)  # before decompose
        assert e.value.code == -5
        with pytest.raises(HgError):
            b.set_params(block=300)
        b.decompose()
        out = np.empty(16, dtype=np.int32)
        assert lib().hg_batch_core_index(b._h, len(gs), out.ctypes.data_as(C.POINTER(C.c_int32))) == -1


def test_batch_rejects_bad_neighbour_id():
    from types import SimpleNamespace as NS
    good = G.graph(4, [[(0, 1), (2, 3)]])
    bad = NS(n=3, L=1, GetN=lambda: 3, layers=lambda: [NS(row_ptr=np.array([0, 1, 2, 2], dtype=np.int64),
                                                         col=np.array([5, 0], dtype=np.uint32))])
    with pytest.raises(HgError) as e:
        Batch([good, bad])
    assert e.value.code == -1


def test_concurrent_batches_from_threads():
    """Batches launched from four host threads at once (their persistent
    kernels share the GPU): every graph still gets the oracle's indices.
    Waves of one workgroup hand FIFO entries and compacted lists to each
    other through global memory; under concurrent kernels a store can still
    be in flight at a plain barrier, which this load exposes."""
    pool = [MultilayerGraph.generate(batch_spec(g, base=0xBD, n_lo=300, n_hi=8000)) for g in range(24)]
    refs = [O.coreindex(mg.n, mg.layers()) for mg in pool]
    say("case concurrent batches from 4 threads")
    bad = []

    def worker(w):
        rng = np.random.default_rng(1000 + w)
        for t in range(30):
            sel = [int(x) for x in rng.integers(0, len(pool), size=int(rng.integers(1, 9)))]
            with Batch([pool[i] for i in sel]) as b:
                b.set_params(block=512, unroll=4, blocks_per_cu=2, order=0, use_graph=0)
                b.decompose()
                for j, i in enumerate(sel):
                    if not np.array_equal(b.core_index(j), refs[i]):
                        bad.append((w, t, i))

    ts = [threading.Thread(target=worker, args=(w,)) for w in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not bad, bad


@pytest.mark.timeout(600)
def test_full_batched_workload_against_oracle(capsys):
    """The full batched workload: 256 graphs (n log-uniform in [1e3, 2e5], L
    in {2..16}, average degree in [2, 20], gamma in [2.1, 3.0]; 744M directed
    entries) in one launch with the default shape and once more through the
    captured-graph replay; every graph's core index and cells equal the
    oracle's CoreIndex restatement (graphs spread over host threads)."""
    import concurrent.futures as cf
    import time

    from pargridcpu_amd.graph import batch_workload

    def note(msg):
        with capsys.disabled():
            print(msg, flush=True)

    t0 = time.perf_counter()
    gs = [MultilayerGraph.generate(s) for s in batch_workload(256)]
    note(f"case batch full workload: 256 graphs, D={sum(g.directed_entries() for g in gs)}, "
         f"generated in {time.perf_counter() - t0:.0f} s")
    with Batch(gs) as b:
        b.decompose()
        first = [b.core_index(i) for i in range(len(gs))]
        cells = [b.cells(i) for i in range(len(gs))]
        b.decompose()
        for i in range(len(gs)):
            assert np.array_equal(b.core_index(i), first[i]), f"replay differs, graph {i}"
    note(f"  batch done at {time.perf_counter() - t0:.0f} s; oracle over 16 host threads")
    order = sorted(range(len(gs)), key=lambda i: -gs[i].directed_entries() * gs[i].L)  # largest first

    def one(i):
        return i, O.coreindex(gs[i].n, gs[i].layers())

    with cf.ThreadPoolExecutor(max_workers=16) as ex:
        for i, ref in ex.map(one, order):
            assert np.array_equal(first[i], ref), f"graph {i}"
            cl, kmax = cells[i]
            assert np.array_equal(kmax, ref.max(axis=1)), f"graph {i}"
            for l in range(gs[i].L):
                h = np.bincount(ref[l], minlength=int(kmax[l]) + 1)
                assert np.array_equal(cl[l, :kmax[l] + 1], np.cumsum(h[::-1])[::-1]), f"graph {i} lmd {l + 1}"
    note(f"  256 graphs exact at {time.perf_counter() - t0:.0f} s")
