"""GPU parity at the billion-edge workload (n = 100M, L = 8, gamma = 2.1,
30 % reshuffle, expected degree capped at 200k, 150M sampled edges per
layer, D = 2.31B directed entries) and at reduced n with the same hub cap.

The engine's result is compared bit-exact, every lmd row, with the oracle's
CoreIndex restatement (get_index2, CoreIndexParallel/CoreIndex.cpp:46-109)
run one lmd per host thread as ParCoreIndex runs it (CoreIndex.cpp:252), and
with the oracle's local FirmCore conditions (valid + tight at every vertex).
Paths that only run at this scale are covered here: hubs of up to ~200k
entries deferred as many 1024-entry chunks, the deferred-list bound, the
per-block lists at the full grid, and core indices of 4096 and above (the
index histogram's global-memory bins).
"""
import dataclasses
import resource
import time

import numpy as np
import pytest

from oracle import oracle as O
from pargridcpu_amd.graph import C3, MultilayerGraph
from pargridcpu_amd.hipgrid import Engine
from tests import graphs as G

pytestmark = pytest.mark.gpu


def c3_sample(n):
    """The 100M-vertex generator at n vertices: the same gamma, reshuffle,
    seed and per-vertex edge density (1.5 n sampled edges per layer) and the
    same 200k expected-degree cap, so the largest hubs still reach the cap."""
    m = int(1.5 * n)
    return dataclasses.replace(C3, name=f"c3-sample-n{n}", n=n, m_lo=m, m_hi=m)


def decompose(mg):
    with Engine.from_graph(mg) as eng:
        eng.decompose()
        ci = eng.core_index()
        cl, kmax = eng.cells()
        t = eng.timing()
    return ci, cl, kmax, t


def check_structure(ci, cl, kmax):
    assert (ci >= 0).all()
    assert (ci[1:] <= ci[:-1]).all()                    # S(k, lmd+1) within S(k, lmd)
    assert (kmax[1:] <= kmax[:-1]).all()                # staircase
    assert np.array_equal(kmax, ci.max(axis=1))
    for i in range(ci.shape[0]):                        # cell lengths = suffix counts
        h = np.bincount(ci[i], minlength=int(kmax[i]) + 1)
        assert np.array_equal(cl[i, :kmax[i] + 1], np.cumsum(h[::-1])[::-1])
        assert (cl[i, kmax[i] + 1:] == 0).all()


def check_oracle(mg, ci, say=print):
    ref = O.coreindex_parallel(mg.n, mg.layers(), threads=min(mg.L, 16),
                               progress=lambda lmd, s: say(f"  oracle lmd={lmd} done at {s:.0f} s"))
    for i in range(mg.L):
        assert np.array_equal(ci[i], ref[i]), f"lmd {i + 1}: {int(np.count_nonzero(ci[i] != ref[i]))} vertices differ"
    res = O.check_local(mg.n, mg.layers(), ci, threads=16)
    assert all(bad == 0 for bad, _ in res.values()), res


@pytest.mark.parametrize("n", [2_000_000, 5_000_000])
def test_c3_sample_against_oracle(n, capsys):
    def say(msg):
        with capsys.disabled():
            print(msg, flush=True)

    mg = MultilayerGraph.generate(c3_sample(n))
    maxdeg = max(int(l.degrees().max()) for l in mg.layers())
    say(f"case c3 sample n={n} D={mg.directed_entries()} max layer degree={maxdeg}")
    ci, cl, kmax, t = decompose(mg)
    check_structure(ci, cl, kmax)
    check_oracle(mg, ci, say)


def test_c3_sample_binned(monkeypatch, capsys):
    """The 2M-vertex sample (46M entries, below the default binning size)
    with the binned apply forced on for every dense first round: exact."""
    monkeypatch.setenv("HG_BIN_MIN", "1")
    mg = MultilayerGraph.generate(c3_sample(2_000_000))
    ci, cl, kmax, t = decompose(mg)
    with capsys.disabled():
        print(f"case c3 sample n=2M binned: {t.binned} binned rounds, {t.decompose_ms:.1f} ms", flush=True)
    assert t.binned > 0
    check_structure(ci, cl, kmax)
    ref = O.coreindex_parallel(mg.n, mg.layers(), threads=8)
    assert np.array_equal(ci, ref)


@pytest.mark.parametrize("heavy,defer", [(256, 256), (1024, 4096)])
def test_c3_sample_deferral_sizes(monkeypatch, heavy, defer, capsys):
    """Hubs far above the deferral threshold split into many chunks of
    another size; exact for both."""
    monkeypatch.setenv("HG_HEAVY", str(heavy))
    monkeypatch.setenv("HG_DEFER", str(defer))
    mg = MultilayerGraph.generate(c3_sample(2_000_000))
    with capsys.disabled():
        print(f"case c3 sample n=2M HG_HEAVY={heavy} HG_DEFER={defer}", flush=True)
    ci, cl, kmax, t = decompose(mg)
    check_structure(ci, cl, kmax)
    ref = O.coreindex_parallel(mg.n, mg.layers(), threads=8)
    assert np.array_equal(ci, ref)


def test_core_index_above_lds_histogram():
    """A 4300-clique in two layers over a sparse random graph: core indices
    of 4299 go through the index histogram's global-memory bins (indices of
    4096 and above); cells and indices match the oracle."""
    n = 6000
    rng = np.random.default_rng(3)
    cl_edges = G.clique(range(4300))
    noise = list(map(tuple, rng.integers(0, n, size=(20000, 2))))
    mg = G.graph(n, [cl_edges + noise, cl_edges + noise[:5000], noise])
    print("case clique 4300 over 6000 vertices, 3 layers", flush=True)
    ci, cl, kmax, t = decompose(mg)
    ref = O.coreindex(mg.n, mg.layers())
    assert np.array_equal(ci, ref)
    assert kmax[0] >= 4299 and kmax[1] >= 4299
    check_structure(ci, cl, kmax)


@pytest.mark.timeout(1000)
def test_c3_full_size(capsys):
    """The 100M-vertex, 2.31B-entry workload at full size: all 8 lmd rows
    bit-exact against the oracle, the local FirmCore conditions at every
    vertex, nesting, staircase and cell lengths.  Prints progress (the
    oracle runs a few minutes on the host)."""
    def say(msg):
        with capsys.disabled():
            print(msg, flush=True)

    t0 = time.perf_counter()
    mg = MultilayerGraph.generate(C3)
    say(f"case c3 full size: generated n={mg.n} L={mg.L} D={mg.directed_entries()} in "
        f"{time.perf_counter() - t0:.0f} s")
    ci, cl, kmax, t = decompose(mg)
    say(f"  engine: {t.decompose_ms:.1f} ms, {t.binned} binned rounds, kmax={kmax.tolist()}")
    assert t.binned > 0  # the default path bins the dense removal rounds of this graph
    check_structure(ci, cl, kmax)
    say(f"  structure ok at {time.perf_counter() - t0:.0f} s; oracle (one lmd per host thread)")
    check_oracle(mg, ci, say)
    rss = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 2**20
    say(f"  c3 full size exact, all lmd; {time.perf_counter() - t0:.0f} s in all, peak RSS {rss:.1f} GiB")
