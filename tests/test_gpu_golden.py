"""GPU tests against the reference's own outputs (tests/golden/) and
size-independent properties at a full benchmark size."""
import dataclasses
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pargridcpu_amd.graph import C2, MultilayerGraph
from pargridcpu_amd.hipgrid import Engine, HipGrid
from tests.golden.make_golden import GRAPHS, LOADER_LAYERS, graph_checksum

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def fixture(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_engine_matches_reference(name):
    fx = fixture(name)
    g = MultilayerGraph.generate(GRAPHS[name])
    assert graph_checksum(g) == str(fx["checksum"])
    say(f"case golden {name} n={g.n} L={g.L}")
    with Engine.from_graph(g) as eng:
        eng.decompose()
        ci = eng.core_index()
        cl, kmax = eng.cells()
        dens = {}
        for beta in (1.0, 0.5, 2.0):
            w = eng.densest(beta)
            dens[beta] = (np.float32(w.max_density), [w.res_k, w.res_lmd, w.res_len], eng.densest_vertices())
    assert np.array_equal(ci, fx["core_index"])
    rows = sorted((i + 1, k, int(cl[i, k])) for i in range(g.L) for k in range(1, int(kmax[i]) + 1))
    assert np.array_equal(np.array(rows, dtype=np.int64).reshape(-1, 3), fx["cells"])
    assert int(kmax.sum()) == int(fx["lpe_count"])
    for beta, (md, res, verts) in dens.items():
        tag = str(beta).replace(".", "p")
        assert md == fx[f"wds_{tag}"][0]
        assert res == fx[f"wds_{tag}_res"].tolist()
        assert np.array_equal(verts, np.nonzero(fx["core_index"][res[1] - 1] >= res[0])[0].astype(np.uint32))


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_wds_coreindex_matches_reference(name):
    """-m WdsCoreIndex: per-lmd records of get_wds and the printed answer."""
    fx = fixture(name + "_wdsci")
    g = MultilayerGraph.generate(GRAPHS[name])
    assert graph_checksum(g) == str(fx["checksum"])
    say(f"case wdscoreindex {name}")
    with Engine.from_graph(g) as eng:
        eng.decompose()
        for beta in (1.0, 0.5, 2.0):
            tag = "ci_" + str(beta).replace(".", "p")
            ans, per = eng.densest_coreindex(beta)
            assert [np.float32(w.max_density) for w in per] == fx[tag].tolist()
            assert [[w.res_k, w.res_lmd] for w in per] == fx[tag + "_res"].tolist()
            ref = O.wds_coreindex(g.n, g.layers(), beta)["answer"]
            assert (np.float32(ans.max_density), ans.res_k, ans.res_lmd) == (ref["max_density"], ref["res_k"],
                                                                              ref["res_lmd"])
            md, rk, rl = fx[tag + "_printed"]
            if g.L > 1:  # with L = 1 the reference prints an uninitialised record
                assert float(f"{float(ans.max_density):g}") == md and (ans.res_k, ans.res_lmd) == (int(rk), int(rl))


def test_loader_case(tmp_path):
    fx = fixture("loader")
    d = tmp_path / "loader"
    d.mkdir()
    (d / "mlg.conf").write_text("".join(f"l{i}.txt\n" for i in range(len(LOADER_LAYERS))))
    for i, text in enumerate(LOADER_LAYERS):
        (d / f"l{i}.txt").write_text(text)
    g = MultilayerGraph.LoadFromFile(str(d) + "/")
    say("case loader")
    tree = HipGrid.Execute(g)
    assert np.array_equal(tree.core_index, fx["core_index"])
    assert tree.getCoreByKAndLmd(2, 1).tolist() == list(range(6))
    assert tree.getCoreByKAndLmd(3, 1) is None


def test_driver_wds_c1(tmp_path):
    fx = fixture("c1")
    g = MultilayerGraph.generate(GRAPHS["c1"])
    root = tmp_path / "dataset"
    g.write_text(str(root / "c1"))
    exe = os.path.join(ROOT, "pargridcpu_amd", "bin", "hipgrid")
    say("case driver HipGridWds c1")
    r = subprocess.run([exe, "-d", "c1", "-m", "HipGridWds", "-o", "1", "-root", str(root)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "|L| = 7, |V| = 20000" in out
    assert re.search(r"count = (\d+)", out).group(1) == str(int(fx["lpe_count"]))
    assert re.search(r"res_k = (\d+)", out).group(1) == str(fx["wds_1p0_res"][0])
    assert re.search(r"res_lmd = (\d+)", out).group(1) == str(fx["wds_1p0_res"][1])
    assert re.search(r"res_len = (\d+)", out).group(1) == str(fx["wds_1p0_res"][2])
    md = float(re.search(r"maximum_density = ([0-9.e+-]+)", out).group(1))
    assert abs(md - float(fx["wds_1p0"][0])) <= 1e-4 * float(fx["wds_1p0"][0])  # printed with %g
    assert "mem = " in out and "HipGrid Elapsed time:" in out


def test_driver_wds_coreindex_c1(tmp_path):
    fx = fixture("c1_wdsci")
    g = MultilayerGraph.generate(GRAPHS["c1"])
    root = tmp_path / "dataset"
    g.write_text(str(root / "c1"))
    exe = os.path.join(ROOT, "pargridcpu_amd", "bin", "hipgrid")
    say("case driver HipGridWdsCoreIndex c1")
    r = subprocess.run([exe, "-d", "c1", "-m", "HipGridWdsCoreIndex", "-root", str(root)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    md, rk, rl = fx["ci_1p0_printed"]
    assert float(re.search(r"maximum_density = (\S+)", r.stdout).group(1)) == md
    assert int(re.search(r"res_k = (\d+)", r.stdout).group(1)) == int(rk)
    assert int(re.search(r"res_lmd = (\d+)", r.stdout).group(1)) == int(rl)
    assert "Wds CoreIndex Time = " in r.stdout


def test_int32_offsets_and_layer_order():
    g = MultilayerGraph.generate(GRAPHS["g2000_L5"])
    fx = fixture("g2000_L5")
    say("case int32 offsets")
    with Engine(g.layers(), g.n, offset_bits=32) as eng:
        eng.decompose()
        assert np.array_equal(eng.core_index(), fx["core_index"])
    say("case layer orders")
    for o in (1, 2, 3, 4):
        g.SetGraphOrder(o)
        with Engine.from_graph(g) as eng:
            eng.decompose()
            assert np.array_equal(eng.core_index(), fx["core_index"])


def test_full_size_c2_properties():
    """C2 at full size (n = 1M, L = 10): bit-exact against the oracle's
    CoreIndex restatement, plus structural invariants and run-to-run
    determinism."""
    g = MultilayerGraph.generate(C2)
    say(f"case C2 full size n={g.n} D={g.directed_entries()}")
    with Engine.from_graph(g) as eng:
        eng.decompose()
        a = eng.core_index()
        cl, kmax = eng.cells()
        off, vtx, kcap = eng.shells()
        eng.decompose()
        b = eng.core_index()
    assert np.array_equal(a, b)
    assert (a[1:] <= a[:-1]).all()                      # nesting in lmd
    assert (kmax[1:] <= kmax[:-1]).all()                # staircase
    assert (a >= 0).all()
    for i in range(g.L):                                # cell lengths = suffix counts
        h = np.bincount(a[i], minlength=int(kmax[i]) + 1)
        assert np.array_equal(cl[i, :kmax[i] + 1], np.cumsum(h[::-1])[::-1])
    stride = kcap + 1
    assert off[-1] == g.n * g.L                         # shells partition every (lmd, v)
    say("case C2 oracle (CPU, ~10 s)")
    ref = O.coreindex(g.n, g.layers())
    assert np.array_equal(a, ref)
    for i in range(g.L):  # every shell: the vertices of each index, ids ascending, k ascending
        h = np.bincount(ref[i], minlength=kcap)
        assert np.array_equal(np.diff(off[i * stride:i * stride + kcap + 1]), h[:kcap])
        seg = vtx[off[i * stride]:off[i * stride + kcap]]
        assert np.array_equal(seg, np.argsort(ref[i], kind="stable").astype(np.uint32))


def _driver_root(tmp_path, dataset):
    """Lays out <tmp>/dataset/<dataset>, <tmp>/run (cwd), <tmp>/output: the
    reference's relative paths (main.cpp:98, Core/FCTree.cpp:141)."""
    for sub in ("dataset", "run", "output"):
        (tmp_path / sub).mkdir(exist_ok=True)
    ds = tmp_path / "dataset" / dataset
    if dataset == "loader":
        ds.mkdir()
        (ds / "mlg.conf").write_text("".join(f"l{i}.txt\n" for i in range(len(LOADER_LAYERS))))
        for i, text in enumerate(LOADER_LAYERS):
            (ds / f"l{i}.txt").write_text(text)
    else:
        MultilayerGraph.generate(GRAPHS[dataset]).write_text(str(ds))
    return tmp_path


@pytest.mark.parametrize("case", ["g2000_L5_k2_lmd2", "g2000_L5_kmax_lmd1", "c1_k3_lmd4", "loader_k2_lmd1"])
def test_driver_query_file_matches_reference(tmp_path, case):
    """`-m HipGrid -k K -lmd L` writes the cell's vertex labels; the same
    header and the same set of labels as the reference driver's
    `-m kpe -k K -lmd L` file (FCTree::saveCoreToLocal, Core/FCTree.cpp:
    138-153, main.cpp:213-223), which lists them in its core[] order."""
    fx = fixture("savecore")
    dataset, k, lmd = str(fx[case + "_dataset"]), int(fx[case + "_k"]), int(fx[case + "_lmd"])
    root = _driver_root(tmp_path, dataset)
    exe = os.path.join(ROOT, "pargridcpu_amd", "bin", "hipgrid")
    say(f"case driver query {case}")
    r = subprocess.run([exe, "-d", dataset, "-m", "HipGrid", "-k", str(k), "-lmd", str(lmd)],
                       cwd=str(root / "run"), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = (root / "output" / f"{dataset}_k_{k}_lmd_{lmd}_HipGrid.txt").read_text().splitlines()
    assert lines[0] == str(fx[case + "_header"])
    got = np.array([int(x) for x in lines[1:]], dtype=np.uint64)
    want = fx[case + "_labels"]
    assert got.size == want.size
    assert np.array_equal(np.sort(got), np.sort(want))
    # a cell above kmax: the reference's "No statisfy result" line
    r = subprocess.run([exe, "-d", dataset, "-m", "HipGrid", "-k", "100000", "-lmd", str(lmd)],
                       cwd=str(root / "run"), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "No statisfy result" in r.stdout


@pytest.mark.parametrize("name", ["g2000_L5", "g4000_L4_loguni"])
def test_driver_mem_shells_match_reference(tmp_path, name):
    """`-m HipGridMem` prints every (k, lmd) shell's size; the reference's
    CoreMem keeps the same shells as father->diff_res
    (CoreParallel/FCSyncMem.cpp:42-51), whose vertices are exactly those of
    golden index k (the fixture script checks CoreMem against lpe)."""
    fx = fixture(name)
    root = _driver_root(tmp_path, name)
    exe = os.path.join(ROOT, "pargridcpu_amd", "bin", "hipgrid")
    say(f"case driver HipGridMem {name}")
    r = subprocess.run([exe, "-d", name, "-m", "HipGridMem"], cwd=str(root / "run"), capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    got = {(int(k), int(l)): int(s) for k, l, s in re.findall(r"shell k = (\d+) lmd = (\d+) size = (\d+)", r.stdout)}
    ci = fx["core_index"]
    want = {(k, i + 1): int(np.count_nonzero(ci[i] == k)) for i in range(ci.shape[0])
            for k in range(1, int(ci[i].max()) + 1)}
    assert got == want


def test_driver_phase_timing(tmp_path):
    """`-phase 1`: the decomposition's level-0, scan and cascade times and
    the busy share, beside the reference's lines; the counts are unchanged."""
    fx = fixture("g2000_L5")
    root = _driver_root(tmp_path, "g2000_L5")
    exe = os.path.join(ROOT, "pargridcpu_amd", "bin", "hipgrid")
    say("case driver phase timing")
    r = subprocess.run([exe, "-d", "g2000_L5", "-m", "HipGrid", "-phase", "1"], cwd=str(root / "run"),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert re.search(r"count = (\d+)", r.stdout).group(1) == str(int(fx["lpe_count"]))
    m = re.search(r"level 0: (\S+) ms, scans: (\S+) ms, cascades: (\S+) ms", r.stdout)
    assert m and all(float(x) >= 0 for x in m.groups())
    u = float(re.search(r"U_overall = (\S+) %", r.stdout).group(1))
    assert 0 < u <= 100.5
