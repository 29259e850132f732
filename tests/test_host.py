"""CPU tests of the host library (libhg_host.so): generator, CSR invariants,
reference text format reader/writer, layer order."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from pargridcpu_amd.graph import C1, C2, MultilayerGraph
from tests.golden.make_golden import GRAPHS, LOADER_LAYERS, graph_checksum

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def check_symmetric_csr(g):
    for lay in g.layers():
        rp, col = lay.row_ptr, lay.col.astype(np.int64)
        assert rp[0] == 0 and (np.diff(rp) >= 0).all() and rp[-1] == col.size
        src = np.repeat(np.arange(g.n), np.diff(rp))
        assert (col < g.n).all()
        assert (src != col).all(), "self-loop"
        key = src * g.n + col
        assert (np.diff(key) > 0).all(), "rows not sorted / duplicates"
        rev = np.sort(col * g.n + src)
        assert np.array_equal(rev, key), "not symmetric"


def test_generator_deterministic_and_valid():
    a = MultilayerGraph.generate(GRAPHS["g2000_L5"])
    b = MultilayerGraph.generate(GRAPHS["g2000_L5"])
    assert graph_checksum(a) == graph_checksum(b)
    check_symmetric_csr(a)
    spec = GRAPHS["g2000_L5"]
    for lay in a.layers():  # at most m sampled pairs survive
        assert lay.getM() <= 2 * spec.m_hi
    import dataclasses

    c = MultilayerGraph.generate(dataclasses.replace(spec, seed=spec.seed + 1))
    assert graph_checksum(c) != graph_checksum(a)


def test_generator_c1_shape():
    g = MultilayerGraph.generate(C1)
    assert g.n == 20000 and g.L == 7
    check_symmetric_csr(g)
    e = [lay.getM() // 2 for lay in g._layers]
    assert all(70000 < x <= 80000 for x in e)  # 80k sampled pairs, minus self-loops and duplicates
    assert g.GetOrder() == sorted(range(7), key=lambda i: (e[i] / 20000, i))  # -o 1: increasing density


def test_generator_loguniform_edge_counts():
    import dataclasses

    spec = dataclasses.replace(C2, n=20000, m_lo=3000, m_hi=30000)
    g = MultilayerGraph.generate(spec)
    m = [lay.getM() // 2 for lay in g._layers]
    assert min(m) >= 0 and max(m) <= 30000 and len(set(m)) > 1


def test_text_roundtrip(tmp_path):
    g = MultilayerGraph.generate(GRAPHS["g300_L3"])
    d = str(tmp_path / "ds")
    g.write_text(d)
    h = MultilayerGraph.LoadFromFile(d + "/")
    labels = h.LoadId2VtxMap().astype(np.int64)
    assert os.path.exists(os.path.join(d, "__vtx_map.txt"))
    # same graph up to the first-appearance relabelling; vertices without
    # edges are not in the files
    present = np.zeros(g.n, dtype=bool)
    for lay in g.layers():
        present |= np.diff(lay.row_ptr) > 0
    assert h.n == int(present.sum())
    assert sorted(labels.tolist()) == np.nonzero(present)[0].tolist()
    for la, lb in zip(g.layers(), h.layers()):
        src_b = np.repeat(np.arange(h.n), np.diff(lb.row_ptr))
        eb = np.sort(labels[src_b] * g.n + labels[lb.col.astype(np.int64)])
        src_a = np.repeat(np.arange(g.n), np.diff(la.row_ptr))
        ea = np.sort(src_a * g.n + la.col.astype(np.int64))
        assert np.array_equal(ea, eb)
    # the core index is invariant under the relabelling
    ci_a = O.coreindex(g.n, g.layers())
    ci_b = O.coreindex(h.n, h.layers())
    assert np.array_equal(ci_a[:, labels], ci_b)


def test_loader_matches_reference(tmp_path):
    fx = dict(np.load(os.path.join(GOLDEN, "loader.npz")))
    d = tmp_path / "loader"
    d.mkdir()
    (d / "mlg.conf").write_text("".join(f"l{i}.txt\n" for i in range(len(LOADER_LAYERS))))
    for i, text in enumerate(LOADER_LAYERS):
        (d / f"l{i}.txt").write_text(text)
    g = MultilayerGraph.LoadFromFile(str(d) + "/")
    assert g.LoadId2VtxMap().tolist() == fx["labels"].tolist()
    assert np.array_equal(O.coreindex(g.n, g.layers()), fx["core_index"])
    vmap = (d / "__vtx_map.txt").read_text().split()
    assert [int(x) for x in vmap[0::2]] == fx["labels"].tolist()
    assert [int(x) for x in vmap[1::2]] == list(range(len(fx["labels"])))


def test_loader_quirks(tmp_path):
    d = tmp_path / "q"
    d.mkdir()
    (d / "mlg.conf").write_text("a.txt\n\nb.txt")  # blank line skipped, no trailing newline
    (d / "a.txt").write_text("1 2\n2 3\n3")        # lone trailing token pairs with the previous v (3 -> 3: dropped)
    (d / "b.txt").write_text("4 1\n1 4")            # no trailing newline
    g = MultilayerGraph.LoadFromFile(str(d) + "/", write_vtx_map=False)
    assert g.L == 2 and g.LoadId2VtxMap().tolist() == [1, 2, 3, 4]
    assert g.GetGraph(0).getM() == 4 and g.GetGraph(1).getM() == 2


@pytest.mark.parametrize("o", [0, 1, 2, 3, 4])
def test_layer_order(o):
    g = MultilayerGraph.generate(GRAPHS["g4000_L4_loguni"])
    g.SetGraphOrder(o)
    m = [lay.getM() for lay in g._layers]
    dens = [np.float32(np.float32(x >> 1) / np.float32(g.n)) for x in m]
    expect = list(range(g.L))
    if o == 1:
        expect.sort(key=lambda i: dens[i])
    elif o == 2:
        expect.sort(key=lambda i: -dens[i])
    elif o == 3:
        expect.sort(key=lambda i: m[i])
    elif o == 4:
        expect.sort(key=lambda i: -m[i])
    assert g.GetOrder() == expect
    # FirmCore is invariant under the layer order
    assert np.array_equal(O.coreindex(g.n, g.layers()), O.coreindex(g.n, g._layers))


def test_statistics_line():
    g = MultilayerGraph.generate(GRAPHS["g300_L3"])
    s = g.PrintStatistics()
    m = [lay.getM() >> 1 for lay in g._layers]
    assert s == f"|L| = 3, |V| = 300, |E| = {sum(m)}, max_|E_i| = {max(m)}"


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asked for more it says so and exits."""
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stdout + r.stderr


def test_committed_profiles_carry_traffic():
    """The bench line's roofline traffic comes from the committed rocprofv3
    summaries; each holds its figure and the commands that produced it."""
    import importlib.util

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(root, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    for name, key in (("hg_decompose_c3", "traffic_bytes_per_decomposition"),
                      ("hg_batch_peel_c4", "traffic_bytes_per_launch")):
        val, src = bench.measured_traffic(name, key)
        assert src == os.path.join("profiles", f"{name}.md")
        assert val is not None and val > 10**10
        text = open(os.path.join(root, src)).read()
        assert "rocprofv3 --kernel-trace --stats" in text and "--pmc FETCH_SIZE" in text and "--pmc WRITE_SIZE" in text


def test_committed_c3_profile_carries_l2_counts():
    """The bench line's L2 hit rate comes from the TCC_HIT/TCC_MISS run in
    the committed summary of the 100M-vertex workload."""
    import importlib.util

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(root, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    hits, src = bench.measured_traffic("hg_decompose_c3", "l2_hits_per_decomposition")
    miss, _ = bench.measured_traffic("hg_decompose_c3", "l2_misses_per_decomposition")
    assert src == os.path.join("profiles", "hg_decompose_c3.md")
    assert hits > 10**9 and miss > 10**9
    assert "--pmc TCC_HIT_sum TCC_MISS_sum" in open(os.path.join(root, src)).read()
