"""CPU tests pinning the oracle (oracle/firmcore_oracle.c).

Each oracle function is checked against
  * the golden vectors written by the reference itself (tests/golden/*.npz,
    produced by tests/golden/make_golden.py from the reference compiled from
    source): core indices, ParGrid cells, WdsList results, lpe's printed
    counters;
  * closed forms: L = 1 is the classic k-core decomposition, identical layers
    give the same index for every lmd, cliques/stars/cycles/paths;
  * a definition-level brute force on small graphs;
  * structural invariants (nesting in k and lmd, staircase, Top-lmd degree
    inside S(k, lmd) >= k, maximality);
  * for WdsList, an independent numpy recomputation of every cell's induced
    edge counts and score.
"""
import glob
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from pargridcpu_amd.graph import MultilayerGraph
from tests import graphs as G
from tests.golden.make_golden import GRAPHS, graph_checksum

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def load_fixture(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))


def fixture_graph(name):
    fx = load_fixture(name)
    g = MultilayerGraph.generate(GRAPHS[name])
    assert graph_checksum(g) == str(fx["checksum"]), "generator output changed: regenerate the golden fixtures"
    return g, fx


def cells_from_index(ci):
    rows = []
    for i in range(ci.shape[0]):
        for k in range(1, int(ci[i].max()) + 1 if ci[i].size else 1):
            rows.append((i + 1, k, int(np.count_nonzero(ci[i] >= k))))
    return np.array(sorted(rows), dtype=np.int64).reshape(-1, 3)


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_oracle_matches_reference(name):
    g, fx = fixture_graph(name)
    lay = g.layers()
    ref = fx["core_index"]
    ci_pg, cl, kmax = O.pargrid(g.n, lay)
    assert np.array_equal(ci_pg, ref)
    assert np.array_equal(cells_from_index(ci_pg), fx["cells"])
    lpe = O.tree(g.n, lay)
    assert np.array_equal(lpe["core_index"], ref)
    assert lpe["count"] == int(fx["lpe_count"])
    efc = fx["lpe_edge_fc"]
    for k, val in efc:
        assert lpe["edge_fc_k1"][k] == val
    kpe = O.tree(g.n, lay, right_spine=True)
    assert np.array_equal(kpe["core_index"], ref)
    assert np.array_equal(O.coreindex(g.n, lay), ref)
    for beta in (1.0, 0.5, 2.0):
        tag = str(beta).replace(".", "p")
        w = O.wds(g.n, lay, beta)
        assert w["max_density"] == fx[f"wds_{tag}"][0]
        assert [w["res_k"], w["res_lmd"], w["res_len"]] == fx[f"wds_{tag}_res"].tolist()


@pytest.mark.parametrize("n,L,m,seed", [(60, 2, 150, 1), (200, 3, 500, 2), (400, 4, 1500, 3), (800, 1, 3000, 4),
                                        (300, 6, 400, 5)])
def test_bruteforce(n, L, m, seed):
    g = G.graph(n, G.random_layers(n, L, m, seed))
    bf = O.bruteforce(n, g.layers())
    assert np.array_equal(O.coreindex(n, g.layers()), bf)
    assert np.array_equal(O.pargrid(n, g.layers())[0], bf)
    assert np.array_equal(O.tree(n, g.layers())["core_index"], bf)


def test_closed_forms():
    n = 40
    es = G.clique(range(6)) + G.star(6, range(7, 15)) + G.cycle(range(15, 25)) + G.path(range(25, 33))
    g = G.graph(n, [es, es, es])
    kc = G.kcore_numbers(n, es)
    for fn in (lambda: O.coreindex(n, g.layers()), lambda: O.pargrid(n, g.layers())[0],
               lambda: O.tree(n, g.layers())["core_index"]):
        ci = fn()
        for i in range(3):  # identical layers: same index for every lmd
            assert np.array_equal(ci[i], kc)
    assert kc[:6].tolist() == [5] * 6 and kc[6:15].tolist() == [1] * 9
    assert kc[15:25].tolist() == [2] * 10 and kc[25:33].tolist() == [1] * 8 and kc[33:].tolist() == [0] * 7
    # L = 1: classic k-core numbers
    rng = np.random.default_rng(9)
    es1 = [tuple(x) for x in rng.integers(0, 500, size=(3000, 2))]
    g1 = G.graph(500, [es1])
    assert np.array_equal(O.coreindex(500, g1.layers())[0], G.kcore_numbers(500, es1))
    # disjoint layers: lmd = 2 needs two layers, so everything is 0
    g2 = G.graph(10, [G.clique(range(5)), G.clique(range(5, 10))])
    ci = O.pargrid(10, g2.layers())[0]
    assert ci[0].tolist() == [4] * 10 and ci[1].tolist() == [0] * 10


def test_invariants_c1():
    g, fx = fixture_graph("c1")
    ci = O.pargrid(g.n, g.layers())[0]
    L = g.L
    # nesting in lmd: S(k, lmd+1) within S(k, lmd)  <=>  idx_{lmd+1} <= idx_lmd
    assert (ci[1:] <= ci[:-1]).all()
    # staircase: kmax non-increasing in lmd
    kmax = ci.max(axis=1)
    assert (kmax[1:] <= kmax[:-1]).all()
    # inside S(k, lmd) every vertex has >= lmd layers with >= k neighbours inside
    lay = g.layers()
    for lmd in range(1, L + 1):
        for k in sorted(set([1, 2, int(kmax[lmd - 1])])):
            if k < 1 or k > kmax[lmd - 1]:
                continue
            inside = ci[lmd - 1] >= k
            good = np.zeros(g.n, dtype=np.int64)
            for lyr in lay:
                src = np.repeat(np.arange(g.n), np.diff(lyr.row_ptr))
                dst = lyr.col.astype(np.int64)
                d = np.bincount(src[inside[src] & inside[dst]], minlength=g.n)
                good += d >= k
            assert (good[inside] >= lmd).all()
    # cell lengths sum to the index sum
    assert sum(int(np.count_nonzero(ci[i] >= k)) for i in range(L) for k in range(1, kmax[i] + 1)) == int(ci.sum())


def wds_numpy(g, ci, beta):
    """Independent restatement of WdsList's scoring on the cells (numpy)."""
    lay = g.layers()
    best = np.float32(0.0)
    res = (0, 0, 0)
    for lmd in range(1, g.L + 1):
        for k in range(1, int(ci[lmd - 1].max()) + 1):
            inside = ci[lmd - 1] >= k
            length = int(inside.sum())
            degs = []
            for lyr in lay:
                src = np.repeat(np.arange(g.n), np.diff(lyr.row_ptr))
                dst = lyr.col.astype(np.int64)
                num_edge = int(np.count_nonzero(inside[src] & inside[dst])) & 0xFFFFFFFF
                degs.append(float(np.float32(np.float32(num_edge) * np.float32(1.0) / np.float32(2))))
            degs.sort(reverse=True)
            dens = 0.0
            for i, dgr in enumerate(degs):
                val = (dgr / length) * math.pow(i + 1, beta)
                if val >= dens:
                    dens = val
            if dens >= float(best):
                best = np.float32(dens)
                res = (k, lmd, length)
    return best, res


@pytest.mark.parametrize("name", ["g300_L3", "g2000_L5", "g1500_L8"])
@pytest.mark.parametrize("beta", [1.0, 0.5, 2.0])
def test_wds_independent(name, beta):
    g, fx = fixture_graph(name)
    ci = fx["core_index"]
    best, res = wds_numpy(g, ci, beta)
    w = O.wds(g.n, g.layers(), beta)
    assert w["max_density"] == best
    assert (w["res_k"], w["res_lmd"], w["res_len"]) == res


@pytest.mark.parametrize("name", sorted(GRAPHS))
@pytest.mark.parametrize("beta", [1.0, 0.5, 2.0])
def test_wds_coreindex_matches_reference(name, beta):
    """WdsCoreIndex restatement against get_wds's per-lmd records and the
    answer CoreIndex::WdsCoreIndex printed (reference built from source)."""
    g, _ = fixture_graph(name)
    fx = load_fixture(name + "_wdsci")
    tag = "ci_" + str(beta).replace(".", "p")
    r = O.wds_coreindex(g.n, g.layers(), beta)
    assert np.array_equal(r["density"], fx[tag])
    assert np.array_equal(np.stack([r["res_k"], r["res_lmd"]], axis=1).astype(np.int64), fx[tag + "_res"])
    md, rk, rl = fx[tag + "_printed"]
    if g.L > 1:  # with L = 1 the reference prints an uninitialised record
        a = r["answer"]
        assert float(f"{float(a['max_density']):g}") == md
        assert (a["res_k"], a["res_lmd"]) == (int(rk), int(rl))


def wds_coreindex_numpy(g, ci, beta):
    """Independent restatement of get_wds's cell scoring (numpy, float32)."""
    best = [(np.float32(0.0), 0, 0) for _ in range(g.L)]
    lay = g.layers()
    fb = np.float32(beta)
    for lmd in range(1, g.L + 1):
        for k in range(1, int(ci[lmd - 1].max()) + 1):
            inside = ci[lmd - 1] >= k
            n_node = int(inside.sum())
            degs = []
            for lyr in lay:
                src = np.repeat(np.arange(g.n), np.diff(lyr.row_ptr))
                cnt = int(np.count_nonzero(inside[src] & inside[lyr.col.astype(np.int64)]))
                degs.append(np.float32(min(cnt, 1 << 24)) / np.float32(2))
            degs.sort(reverse=True)
            for i, d in enumerate(degs):
                mad = np.float32(d / np.float32(n_node))
                dens = np.float32(float(mad) * math.pow(i + 1, float(fb)))
                if dens >= best[lmd - 1][0]:
                    best[lmd - 1] = (dens, k, lmd)
    return best


@pytest.mark.parametrize("name", ["g300_L3", "g2000_L5", "g1500_L8"])
def test_wds_coreindex_independent(name):
    g, fx = fixture_graph(name)
    r = O.wds_coreindex(g.n, g.layers(), 0.5)
    for i, (d, k, l) in enumerate(wds_coreindex_numpy(g, fx["core_index"], 0.5)):
        assert (r["density"][i], r["res_k"][i], r["res_lmd"][i]) == (d, k, l)


def test_float_counter_saturates_at_2_pow_24():
    """A float incremented by 1 is exact up to 2^24 and stays there: the
    reference's per-layer edge counter equals min(count, 2^24)."""
    f = np.float32(1 << 24) - np.float32(3)
    seq = []
    for _ in range(8):
        f = np.float32(f + np.float32(1))
        seq.append(int(f))
    assert seq == [16777214, 16777215, 16777216, 16777216, 16777216, 16777216, 16777216, 16777216]


def test_planted_mistake_is_caught():
    """A mistake planted in a copy of the oracle's result (a dropped vertex
    in one cell) must fail the reference comparison."""
    g, fx = fixture_graph("g2000_L5")
    ci = O.coreindex(g.n, g.layers()).copy()
    v = int(np.argmax(ci[0]))
    ci[0, v] -= 1
    assert not np.array_equal(ci, fx["core_index"])
    assert not np.array_equal(cells_from_index(ci), fx["cells"])


def test_fixture_set_complete():
    names = {os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz"))}
    assert set(GRAPHS) | {"loader"} <= names


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_coreindex_parallel_matches_reference(name):
    """The per-lmd entry run from concurrent threads (ParCoreIndex's shape)
    gives the reference's golden index."""
    g, fx = fixture_graph(name)
    assert np.array_equal(O.coreindex_parallel(g.n, g.layers(), threads=4), fx["core_index"])
    lmd = g.L
    assert np.array_equal(O.coreindex_lmd(g.n, g.layers(), lmd), fx["core_index"][lmd - 1])


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_local_check_accepts_reference_and_rejects_off_by_one(name):
    """orc_check_lmd passes the reference's own index and flags a planted
    +1 / -1 at single vertices (the mistake is planted in a copy of the
    golden index, on the CPU)."""
    g, fx = fixture_graph(name)
    ref = fx["core_index"]
    res = O.check_local(g.n, g.layers(), ref, threads=3)
    assert all(bad == 0 for bad, _ in res.values()), res
    rng = np.random.default_rng(7)
    for lmd in sorted({1, (g.L + 1) // 2, g.L}):
        row = ref[lmd - 1]
        cands = np.flatnonzero(row > 0)
        for v in rng.choice(cands, size=min(4, cands.size), replace=False):
            for delta in (+1, -1):
                bad = ref.copy()
                bad[lmd - 1, v] += delta
                r = O.check_local(g.n, g.layers(), bad, threads=2, lmds=[lmd])[lmd]
                assert r[0] > 0, (lmd, int(v), delta)
        zeros = np.flatnonzero(row == 0)
        if zeros.size:
            bad = ref.copy()
            bad[lmd - 1, zeros[0]] = 1
            assert O.check_local(g.n, g.layers(), bad, lmds=[lmd])[lmd][0] > 0


def test_savecore_fixture_is_the_golden_cell():
    """The reference driver's saveCoreToLocal files (savecore.npz) hold
    exactly S(k, lmd) of the golden index, by label, with the length in the
    header."""
    fx = load_fixture("savecore")
    cases = sorted({k[:-len("_header")] for k in fx if k.endswith("_header")})
    assert len(cases) == 4
    for case in cases:
        dataset, k, lmd = str(fx[case + "_dataset"]), int(fx[case + "_k"]), int(fx[case + "_lmd"])
        labels = fx[case + "_labels"]
        assert str(fx[case + "_header"]) == f"########## k_{k}_lmd_{lmd}_length_{labels.size}"
        if dataset == "loader":
            lf = load_fixture("loader")
            ids = np.flatnonzero(lf["core_index"][lmd - 1] >= k)
            want = lf["labels"][ids]
        else:
            want = np.flatnonzero(load_fixture(dataset)["core_index"][lmd - 1] >= k).astype(np.uint64)
        assert np.array_equal(np.sort(labels), np.sort(want)), case
