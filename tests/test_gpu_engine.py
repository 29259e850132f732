"""GPU parity tests: libhipgrid.so through its C ABI against the oracle.

Bar: bit-exact core indices, cell lengths, k-shells and WdsList results.
Full-size workloads are checked through size-independent properties in
test_gpu_golden.py (test_full_size_c2_properties); here the oracle finishes in seconds.
"""
import dataclasses
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pargridcpu_amd.graph import C1, MultilayerGraph
from pargridcpu_amd.hipgrid import Engine, HgError, HipGrid
from tests import graphs as G

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def small_spec(n, L, m, seed, gamma=2.3, rho=0.5):
    return dataclasses.replace(C1, name=f"t{n}-{L}-{m}-{seed}", seed=seed, n=n, L=L, m_lo=m, m_hi=m, gamma=gamma,
                               rho=rho, order=0)


def run_engine(mg):
    with Engine.from_graph(mg) as eng:
        eng.decompose()
        ci = eng.core_index()
        cl, kmax = eng.cells()
        t = eng.timing()
    return ci, cl, kmax, t


def check_against_oracle(mg, ci, cl, kmax):
    ref = O.coreindex(mg.n, mg.layers())
    assert np.array_equal(ci, ref)
    assert np.array_equal(kmax, ref.max(axis=1))
    for i in range(mg.L):
        for k in range(cl.shape[1]):
            assert cl[i, k] == (np.count_nonzero(ref[i] >= k) if k <= kmax[i] else 0)


@pytest.mark.parametrize("n,L,m,seed", [(50, 1, 120, 1), (200, 3, 600, 2), (1000, 4, 3000, 3), (3000, 7, 8000, 4),
                                        (2000, 12, 2500, 5), (500, 2, 20000, 6)])
def test_seeded_small(n, L, m, seed):
    mg = MultilayerGraph.generate(small_spec(n, L, m, seed))
    say(f"case seeded n={n} L={L} m={m} seed={seed}")
    ci, cl, kmax, t = run_engine(mg)
    check_against_oracle(mg, ci, cl, kmax)
    # ParGrid restatement agrees as well (cells from the reference flow)
    pci, pcl, pk = O.pargrid(mg.n, mg.layers(), cap=cl.shape[1])
    assert np.array_equal(ci, pci)
    assert np.array_equal(cl, pcl)


@pytest.mark.parametrize("heavy,blocks", [(4, 0), (4, 5), (64, 3)])
def test_deferred_chunks(monkeypatch, heavy, blocks):
    """With a tiny heavy threshold almost every removal is deferred to the
    next step as chunks (a small grid gives each block many chunks of every
    instance); the results stay exact, also on a second decomposition of the
    same graph."""
    monkeypatch.setenv("HG_HEAVY", str(heavy))
    if blocks:
        monkeypatch.setenv("HG_BLOCKS", str(blocks))
    for n, L, m, seed in [(3000, 5, 12000, 11), (1500, 8, 9000, 12), (800, 3, 30000, 13)]:
        mg = MultilayerGraph.generate(small_spec(n, L, m, seed))
        say(f"case deferred heavy={heavy} blocks={blocks} n={n} L={L} m={m} seed={seed}")
        ref = O.coreindex(mg.n, mg.layers())
        with Engine.from_graph(mg) as eng:
            for rep in range(2):
                eng.decompose()
                assert np.array_equal(eng.core_index(), ref), f"repeat {rep}"


def test_c1_workload():
    mg = MultilayerGraph.generate(C1)
    say("case C1 n=20000 L=7")
    ci, cl, kmax, t = run_engine(mg)
    check_against_oracle(mg, ci, cl, kmax)
    assert t.levels >= int(kmax.max()) + 1


def test_closed_forms():
    n = 40
    # identical layers of a 6-clique + star + cycle + path: index = k-core number for every lmd
    es = G.clique(range(6)) + G.star(6, range(7, 15)) + G.cycle(range(15, 25)) + G.path(range(25, 33))
    mg = G.graph(n, [es, es, es])
    say("case closed forms")
    ci, cl, kmax, _ = run_engine(mg)
    kc = G.kcore_numbers(n, es)
    for i in range(3):
        assert np.array_equal(ci[i], kc)
    assert list(ci[0][:6]) == [5] * 6
    assert ci[0][6] == 1 and all(ci[0][7:15] == 1)
    assert all(ci[0][15:25] == 2) and all(ci[0][25:33] == 1) and all(ci[0][33:] == 0)
    # L = 1 is the classic k-core decomposition
    rng = np.random.default_rng(7)
    es1 = [tuple(x) for x in rng.integers(0, 300, size=(1500, 2))]
    mg1 = G.graph(300, [es1])
    ci1, _, _, _ = run_engine(mg1)
    assert np.array_equal(ci1[0], G.kcore_numbers(300, es1))


def test_edge_cases():
    say("case no edges")
    mg = G.graph(10, [[], []])
    ci, cl, kmax, _ = run_engine(mg)
    assert (ci == 0).all() and (kmax == 0).all()
    say("case single edge, one layer of two")
    mg = G.graph(3, [[(0, 1)], []])
    ci, cl, kmax, _ = run_engine(mg)
    assert ci.tolist() == [[1, 1, 0], [0, 0, 0]]
    say("case one vertex")
    mg = G.graph(1, [[]])
    ci, _, _, _ = run_engine(mg)
    assert ci.tolist() == [[0]]
    say("case hub (chunk splitting)")
    hub_edges = G.star(0, range(1, 5000)) + G.cycle(range(1, 5000))
    mg = G.graph(5000, [hub_edges, hub_edges[::2]])
    ci, cl, kmax, _ = run_engine(mg)
    check_against_oracle(mg, ci, cl, kmax)


def test_repeat_and_reuse():
    mg = MultilayerGraph.generate(small_spec(4000, 5, 12000, 11))
    say("case repeat decompose x3")
    with Engine.from_graph(mg) as eng:
        outs = []
        for _ in range(3):
            eng.decompose()
            outs.append(eng.core_index())
    assert all(np.array_equal(outs[0], o) for o in outs[1:])
    check_against_oracle(mg, outs[0], *Engine_cells(mg))


def Engine_cells(mg):
    with Engine.from_graph(mg) as eng:
        eng.decompose()
        return eng.cells()


def test_shells():
    """Every (k, lmd) shell, FCSyncMem's diff_res (CoreParallel/FCSyncMem.cpp:
    42-51), against the oracle's core index: the vertices of index k,
    ascending."""
    mg = MultilayerGraph.generate(small_spec(2500, 4, 7000, 12))
    say("case shells")
    with Engine.from_graph(mg) as eng:
        eng.decompose()
        off, vtx, kcap = eng.shells()
    ref = O.coreindex(mg.n, mg.layers())
    stride = kcap + 1
    assert off[-1] == mg.n * mg.L
    for i in range(mg.L):
        for k in range(kcap):
            a, b = off[i * stride + k], off[i * stride + k + 1]
            expect = np.nonzero(ref[i] == k)[0]
            assert np.array_equal(vtx[a:b], expect.astype(np.uint32))


@pytest.mark.parametrize("beta", [1.0, 0.5, 2.0])
def test_densest(beta):
    mg = MultilayerGraph.generate(small_spec(3000, 6, 9000, 13))
    say(f"case densest beta={beta}")
    ref = O.wds(mg.n, mg.layers(), beta)
    with Engine.from_graph(mg) as eng:
        eng.decompose()
        w = eng.densest(beta)
        cap = ref["cell_density"].shape[1]
        dens, sel = eng.cell_density(cap)
        verts = eng.densest_vertices()
        ci = eng.core_index()
    assert np.float32(w.max_density) == ref["max_density"]
    assert (w.res_k, w.res_lmd, w.res_len) == (ref["res_k"], ref["res_lmd"], ref["res_len"])
    assert np.array_equal(dens, ref["cell_density"])
    assert np.array_equal(sel, ref["cell_sel"])
    assert np.array_equal(verts, np.nonzero(ci[w.res_lmd - 1] >= w.res_k)[0].astype(np.uint32))
    assert verts.size == w.res_len


def test_densest_c1():
    mg = MultilayerGraph.generate(C1)
    say("case densest C1")
    ref = O.wds(mg.n, mg.layers(), 1.0)
    m, k, lmd, ln, verts = HipGrid.WeightDenestSubgraph(mg, 1.0)
    assert (m, k, lmd, ln) == (ref["max_density"], ref["res_k"], ref["res_lmd"], ref["res_len"])


def test_errors():
    say("case errors")
    mg = G.graph(4, [[(0, 1)]])
    with pytest.raises(HgError) as e:
        Engine(mg.layers(), 4, offset_bits=16)
    assert e.value.code == -1
    with Engine.from_graph(mg) as eng:
        with pytest.raises(HgError) as e:
            eng.core_index()
        assert e.value.code == -5
        with pytest.raises(HgError):
            eng.densest(1.0)


def test_invalid_graphs_rejected():
    """Offsets that decrease and neighbour ids >= n are caught by the device
    set-up and reported as argument errors, not decomposed."""
    say("case invalid graphs")
    from types import SimpleNamespace as NS
    good = NS(row_ptr=np.array([0, 1, 2, 2], dtype=np.int64), col=np.array([1, 0], dtype=np.uint32))
    down = NS(row_ptr=np.array([0, 2, 1, 2], dtype=np.int64), col=np.array([1, 0], dtype=np.uint32))
    wide = NS(row_ptr=np.array([0, 1, 2, 2], dtype=np.int64), col=np.array([7, 0], dtype=np.uint32))
    with Engine([good], 3) as eng:
        eng.decompose()
        assert eng.core_index().tolist() == [[1, 1, 0]]
    for bad in (down, wide):
        with pytest.raises(HgError) as e:
            Engine([good, bad], 3)
        assert e.value.code == -1


@pytest.mark.parametrize("seed", [11, 12])
def test_int32_offsets(seed):
    """32-bit caller offsets (widened on the device) give the same results."""
    mg = MultilayerGraph.generate(small_spec(4000, 5, 9000, seed))
    say(f"case int32 offsets seed={seed}")
    with Engine(mg.layers(), mg.n, offset_bits=32) as eng:
        eng.decompose()
        ci = eng.core_index()
    assert np.array_equal(ci, O.coreindex(mg.n, mg.layers()))


def test_wds_coreindex_float_counter_saturation():
    """A 4200-clique in two identical layers: every cell is the whole clique
    with 4200 * 4199 > 2^24 directed entries per layer, so the reference's
    float edge counter stops at 2^24 (CoreIndexParallel/CoreIndex.cpp:317-324)."""
    n = 4200
    es = G.clique(range(n))
    mg = G.graph(n, [es, es])
    say("case wdscoreindex saturation (4200-clique, 2 layers)")
    with Engine.from_graph(mg) as eng:
        eng.decompose()
        ci = eng.core_index()
        assert (ci == n - 1).all()
        ans, per = eng.densest_coreindex(1.0)
    mad = np.float32(np.float32(1 << 24) / np.float32(2)) / np.float32(n)
    want = np.float32(float(mad) * 2.0)  # the second-ranked layer, pow(2, 1)
    assert [np.float32(w.max_density) for w in per] == [want, want]
    assert [(w.res_k, w.res_lmd) for w in per] == [(n - 1, 1), (n - 1, 2)]
    assert (np.float32(ans.max_density), ans.res_k, ans.res_lmd) == (want, n - 1, 1)  # lmd = L takes no part



def test_phase_timing():
    """With phase timing on, the scan and cascade sums are measured and add
    up (with level 0) to about the whole decomposition; off, they are -1;
    results are the same either way."""
    mg = MultilayerGraph.generate(small_spec(20000, 6, 60000, 21))
    say("case phase timing n=20000 L=6")
    ref = O.coreindex(mg.n, mg.layers())
    with Engine.from_graph(mg) as eng:
        eng.decompose()
        t = eng.timing()
        assert t.scan_ms == -1 and t.cascade_ms == -1
        eng.set_phase_timing(True)
        eng.decompose()
        t = eng.timing()
        assert np.array_equal(eng.core_index(), ref)
        assert t.scan_ms > 0 and t.cascade_ms > 0 and t.init_ms > 0
        parts = t.init_ms + t.scan_ms + t.cascade_ms
        assert parts <= t.decompose_ms * 1.001 + 0.01
        assert parts >= 0.5 * t.decompose_ms
        eng.set_phase_timing(False)
        eng.decompose()
        assert eng.timing().scan_ms == -1


@pytest.mark.parametrize("n,L,m,seed,gamma", [(3000, 5, 12000, 31, 2.3), (20000, 8, 90000, 32, 2.1),
                                              (60000, 4, 400000, 33, 2.2), (500, 2, 20000, 34, 2.5)])
def test_row_compaction(monkeypatch, n, L, m, seed, gamma):
    """HG_COMPACT=2 compacts rows between levels on every graph (off by
    default): exact against the oracle, also on a second decomposition of the
    same handle (which starts from the full rows again) and with a small grid
    (many chunks per block)."""
    monkeypatch.setenv("HG_COMPACT", "2")
    mg = MultilayerGraph.generate(small_spec(n, L, m, seed, gamma=gamma))
    ref = O.coreindex(mg.n, mg.layers())
    for blocks in (0, 7):
        if blocks:
            monkeypatch.setenv("HG_BLOCKS", str(blocks))
        say(f"case compaction n={n} L={L} m={m} seed={seed} blocks={blocks}")
        with Engine.from_graph(mg) as eng:
            for rep in range(2):
                eng.decompose()
                assert np.array_equal(eng.core_index(), ref), f"repeat {rep}"
                cl, kmax = eng.cells()
                assert np.array_equal(kmax, ref.max(axis=1))


def force_binning(monkeypatch, coarse=None):
    """Every scan step with a non-empty frontier takes the binned apply
    (HG_BIN_MIN=1: any graph; HG_BIN_MINENT=1, HG_BIN_DENS=0: any density)."""
    monkeypatch.setenv("HG_BIN_MIN", "1")
    monkeypatch.setenv("HG_BIN_MINENT", "1")
    monkeypatch.setenv("HG_BIN_DENS", "0")
    monkeypatch.setenv("HG_BIN_STEPS", "100000")
    if coarse:
        monkeypatch.setenv("HG_BIN_C", str(coarse))


@pytest.mark.parametrize("n,L,m,seed,gamma,coarse", [(3000, 5, 12000, 41, 2.3, None), (20000, 8, 90000, 42, 2.1, 64),
                                                     (60000, 4, 400000, 43, 2.2, 1), (500, 2, 20000, 44, 2.5, 8),
                                                     (2000, 12, 2500, 45, 2.3, None), (200, 1, 600, 46, 2.3, 3),
                                                     (300000, 32, 40000, 47, 2.2, 2)])
def test_binned_first_rounds(monkeypatch, n, L, m, seed, gamma, coarse):
    """The binned apply of a level's first round (records grouped by coarse
    and fine target range, applied in LDS, touched rows recounted) forced on
    every scan step: exact against the oracle, also on a repeat of the same
    handle and with a small grid.  Coarse bins from one (HG_BIN_C=1: ranges
    split only by the fine-bin limit; the 32-layer case has 262144 vertices
    per coarse bin at most, so n = 300000 splits) to hundreds."""
    force_binning(monkeypatch, coarse)
    mg = MultilayerGraph.generate(small_spec(n, L, m, seed, gamma=gamma))
    ref = O.coreindex(mg.n, mg.layers())
    for blocks in (0, 5):
        if blocks:
            monkeypatch.setenv("HG_BLOCKS", str(blocks))
        say(f"case binned n={n} L={L} m={m} seed={seed} coarse={coarse} blocks={blocks}")
        with Engine.from_graph(mg) as eng:
            for rep in range(2):
                eng.decompose()
                assert np.array_equal(eng.core_index(), ref), f"repeat {rep}"
                cl, kmax = eng.cells()
                assert np.array_equal(kmax, ref.max(axis=1))
                t = eng.timing()
                if int(ref.max()) >= 1:
                    assert t.binned > 0


def test_binned_with_deferral_and_compaction(monkeypatch):
    """The binned apply beside tiny deferral thresholds (its removals become
    many chunks) and beside row compaction (it reads the compacted rows)."""
    force_binning(monkeypatch, 16)
    mg = MultilayerGraph.generate(small_spec(20000, 6, 80000, 48, gamma=2.1))
    ref = O.coreindex(mg.n, mg.layers())
    for env in ({"HG_HEAVY": "4", "HG_DEFER": "256"}, {"HG_COMPACT": "2"}):
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        say(f"case binned with {env}")
        with Engine.from_graph(mg) as eng:
            eng.decompose()
            assert np.array_equal(eng.core_index(), ref)
            assert eng.timing().binned > 0
        for k in env:
            monkeypatch.delenv(k)


@pytest.mark.parametrize("rounds", [1, 8])
def test_binned_rounds(monkeypatch, rounds):
    """Removal rounds after the first are binned too (round r takes the
    vertices round r - 1 removed) up to HG_BIN_ROUNDS per level; one round
    only, or up to eight: the same result."""
    force_binning(monkeypatch, 32)
    monkeypatch.setenv("HG_BIN_ROUNDS", str(rounds))
    mg = MultilayerGraph.generate(small_spec(20000, 6, 80000, 50, gamma=2.1))
    ref = O.coreindex(mg.n, mg.layers())
    say(f"case binned rounds={rounds}")
    with Engine.from_graph(mg) as eng:
        for rep in range(2):
            eng.decompose()
            assert np.array_equal(eng.core_index(), ref), f"repeat {rep}"
            assert eng.timing().binned > 0


@pytest.mark.parametrize("hbound,binned", [("1", False), ("0", False), ("1", True), ("0", True)])
def test_core_bounds_dense(monkeypatch, hbound, binned):
    """Per-entry core-index bounds (the lmd-th largest layer degree of the
    target, saturated at 255) skip targets that left at an earlier level
    without probing their liveness bit.  A dense graph whose core indices pass
    255 for every lmd exercises the saturated bytes at levels above 255;
    HG_HBOUND=0 keeps the z pre-filter only.  Cascade only and with every scan
    step binned: exact against the oracle."""
    monkeypatch.setenv("HG_HBOUND", hbound)
    if binned:
        force_binning(monkeypatch)
    mg = MultilayerGraph.generate(small_spec(1500, 3, 500000, 54, gamma=3.0))
    say(f"case dense bounds HG_HBOUND={hbound} binned={binned}")
    ci, cl, kmax, t = run_engine(mg)
    assert kmax.min() > 255
    check_against_oracle(mg, ci, cl, kmax)
    if binned:
        assert t.binned > 0


def test_binned_off_matches(monkeypatch):
    """HG_BIN=2 keeps every round on the atomic path (no binned rounds), with
    the same result."""
    force_binning(monkeypatch)
    monkeypatch.setenv("HG_BIN", "2")
    mg = MultilayerGraph.generate(small_spec(3000, 5, 12000, 49))
    with Engine.from_graph(mg) as eng:
        eng.decompose()
        assert eng.timing().binned == 0
        assert np.array_equal(eng.core_index(), O.coreindex(mg.n, mg.layers()))
