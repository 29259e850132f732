#!/usr/bin/env python3
"""Writes the golden fixtures in tests/golden/ from the reference itself.

Runs in the build container only (the reference sources are not shipped):
the reference's own routines are compiled from its sources by
oracle/Makefile (`make -C oracle ref`) into oracle/_ref/ref_harness, which
this script drives on seeded synthetic graphs written in the reference's
text format.  Only inputs' identities (generator parameters + a checksum of
the CSR) and the reference's outputs are stored:

  core_index   int32 [L][n]   per-vertex FirmCore index (from lpe's cell
                              lists; kpe, CoreMem shells and CoreIndex must
                              agree or the script stops)
  cells        int64 [c][3]   (lmd, k, length) of every non-empty cell from
                              ParGrid (FCSyncLeft::Execute); ParGridPlus and
                              ParGridright must agree
  wds_<beta>   WdsList result: float32 max density, res_k, res_lmd, res_len
  lpe_count, lpe_edge_fc      lpe's printed `count` and num_of_edge_fc lines
  labels       uint64 [n_ref] the reference's id -> label map (loader case)

and, per graph, <name>_wdsci.npz from CoreIndex::WdsCoreIndex
(CoreIndexParallel/CoreIndex.cpp:307-478):
  ci_<beta>          float32 [L] get_wds's density per lmd (record values)
  ci_<beta>_res      int64 [L][2] its res_k, res_lmd
  ci_<beta>_printed  the printed answer: density (6 significant digits),
                     res_k, res_lmd

and savecore.npz from the reference's own driver (oracle/_ref/firmcore_ref,
main.cpp) run as `-m kpe -k K -lmd L`, which writes the cell's vertex
labels through FCTree::saveCoreToLocal (Core/FCTree.cpp:138-153,
main.cpp:213-223):
  <case>_header   the file's first line
  <case>_labels   uint64, the labels in the file's order
  <case>_k, <case>_lmd, <case>_dataset

Usage: python tests/golden/make_golden.py [--wdsci-only | --savecore-only]

Run after __graft_entry__.build().
"""
from __future__ import annotations

import dataclasses
import hashlib
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from pargridcpu_amd.graph import C1, MultilayerGraph  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "firmcore_ref")
BETAS = (1.0, 0.5, 2.0)

# name -> generator spec (seeded; see include/hg_host.h)
GRAPHS = {
    "g300_L3": dataclasses.replace(C1, name="g300_L3", seed=101, n=300, L=3, m_lo=900, m_hi=900, gamma=2.3, order=0),
    "g2000_L5": dataclasses.replace(C1, name="g2000_L5", seed=102, n=2000, L=5, m_lo=5000, m_hi=5000, gamma=2.3,
                                    order=0),
    "g1000_L1": dataclasses.replace(C1, name="g1000_L1", seed=103, n=1000, L=1, m_lo=4000, m_hi=4000, gamma=2.2,
                                    order=0),
    "g1500_L8": dataclasses.replace(C1, name="g1500_L8", seed=104, n=1500, L=8, m_lo=2000, m_hi=2000, gamma=2.6,
                                    rho=0.3, order=0),
    "g4000_L4_loguni": dataclasses.replace(C1, name="g4000_L4_loguni", seed=105, n=4000, L=4, m_lo=2000,
                                           m_hi=20000, gamma=2.2, order=0),
    "c1": dataclasses.replace(C1, order=0),
}


def graph_checksum(g: MultilayerGraph) -> str:
    h = hashlib.sha256()
    h.update(np.int64(g.n).tobytes())
    for lay in g.layers():
        h.update(np.ascontiguousarray(lay.row_ptr, dtype=np.int64).tobytes())
        h.update(np.ascontiguousarray(lay.col, dtype=np.uint32).tobytes())
    return h.hexdigest()


def run_harness(mode: str, dataset_dir: str, prefix: str, beta: float = 1.0) -> str:
    r = subprocess.run([HARNESS, mode, dataset_dir + "/", prefix, repr(beta)], capture_output=True, text=True,
                       timeout=3600)
    if r.returncode != 0:
        raise RuntimeError(f"harness {mode} failed: {r.stderr[-2000:]}")
    return r.stdout


def ref_index(prefix: str, n: int, L: int) -> np.ndarray:
    labels = np.fromfile(prefix + ".labels", dtype=np.uint64)
    idx = np.fromfile(prefix + ".index", dtype=np.int32).reshape(L, labels.size)
    full = np.zeros((L, n), dtype=np.int32)
    full[:, labels.astype(np.int64)] = idx  # vertices absent from the files have index 0
    return full


def ref_cells(prefix: str) -> np.ndarray:
    rows = []
    with open(prefix + ".cells") as f:
        for line in f:
            p = line.split()
            rows.append((int(p[0]), int(p[1]), int(p[2])))
    return np.array(sorted(rows), dtype=np.int64).reshape(-1, 3)


def make_graph_fixture(name: str, spec, workdir: str) -> dict:
    g = MultilayerGraph.generate(spec)
    n, L = g.n, g.L
    ds = os.path.join(workdir, name)
    g.write_text(ds)
    out = {"name": name, "seed": spec.seed, "n": n, "L": L, "gamma": spec.gamma, "rho": spec.rho, "m_lo": spec.m_lo,
           "m_hi": spec.m_hi, "cap_deg": spec.cap_deg, "checksum": graph_checksum(g)}
    stdout = run_harness("lpe", ds, os.path.join(workdir, name + "_lpe"))
    lpe = ref_index(os.path.join(workdir, name + "_lpe"), n, L)
    m = re.search(r"^count = (\d+)", stdout, re.M)
    out["lpe_count"] = int(m.group(1))
    efc = [(int(a), int(b)) for a, b in re.findall(r"k = (\d+), lmd = 1, num_of_edge_fc = (\d+)", stdout)]
    out["lpe_edge_fc"] = np.array(efc, dtype=np.int64).reshape(-1, 2)
    for mode in ("kpe", "coremem", "coreindex"):
        run_harness(mode, ds, os.path.join(workdir, f"{name}_{mode}"))
        other = ref_index(os.path.join(workdir, f"{name}_{mode}"), n, L)
        if not np.array_equal(other, lpe):
            raise RuntimeError(f"{name}: reference {mode} disagrees with lpe")
    out["core_index"] = lpe
    cells = None
    for mode in ("pargrid", "pargridplus", "pargridright"):
        run_harness(mode, ds, os.path.join(workdir, f"{name}_{mode}"))
        c = ref_cells(os.path.join(workdir, f"{name}_{mode}"))
        if cells is None:
            cells = c
        elif not np.array_equal(c, cells):
            raise RuntimeError(f"{name}: reference {mode} cells disagree with pargrid")
    out["cells"] = cells
    for beta in BETAS:
        pre = os.path.join(workdir, f"{name}_wds_{beta}")
        run_harness("wds", ds, pre, beta)
        with open(pre + ".wds") as f:
            md, rk, rl, rn = f.read().split()
        tag = str(beta).replace(".", "p")
        out[f"wds_{tag}"] = np.array([float(md)], dtype=np.float32)
        out[f"wds_{tag}_res"] = np.array([int(rk), int(rl), int(rn)], dtype=np.int64)
    return out


# A hand-written edge-list case for the reference loader's id assignment and
# cleaning (Graph/MultilayerGraph.cpp:104-166): labels are first-appearance
# ordered across layers, self-loops dropped (a vertex seen only in a
# self-loop does not exist), duplicates and reversed duplicates merged,
# trailing newline.
LOADER_LAYERS = [
    "10 20\n20 30\n30 10\n10 20\n20 10\n40 40\n30 50\n",
    "50 60\n60 70\n70 50\n99 99\n10 70\n70 10\n",
    "20 30\n30 60\n60 20\n20 30\n",
]


def make_loader_fixture(workdir: str) -> dict:
    ds = os.path.join(workdir, "loader")
    write_loader_dataset(ds)
    pre = os.path.join(workdir, "loader_lpe")
    run_harness("lpe", ds, pre)
    labels = np.fromfile(pre + ".labels", dtype=np.uint64)
    idx = np.fromfile(pre + ".index", dtype=np.int32).reshape(len(LOADER_LAYERS), labels.size)
    return {"name": "loader", "layers": np.array(LOADER_LAYERS, dtype=object).astype(str), "labels": labels,
            "core_index": idx}


def make_wdsci_fixture(name: str, spec, workdir: str) -> dict:
    g = MultilayerGraph.generate(spec)
    ds = os.path.join(workdir, name)
    if not os.path.exists(ds):
        g.write_text(ds)
    out = {"name": name, "checksum": graph_checksum(g)}
    for beta in BETAS:
        pre = os.path.join(workdir, f"{name}_wdsci_{beta}")
        stdout = run_harness("wdscoreindex", ds, pre, beta)
        rows = [line.split() for line in open(pre + ".wdsci")]
        tag = "ci_" + str(beta).replace(".", "p")
        out[tag] = np.array([float(r[1]) for r in rows], dtype=np.float32)
        out[tag + "_res"] = np.array([[int(r[2]), int(r[3])] for r in rows], dtype=np.int64)
        md = float(re.search(r"^maximum_density = (\S+)", stdout, re.M).group(1))
        rk = int(re.search(r"^res_k = (\d+)", stdout, re.M).group(1))
        rl = int(re.search(r"^res_lmd = (\d+)", stdout, re.M).group(1))
        out[tag + "_printed"] = np.array([md, rk, rl], dtype=np.float64)
    return out


# (case, dataset, k, lmd) for savecore.npz; "loader" is the hand-written
# edge-list case below, the others are generated graphs written in the
# reference's text format (labels = generator ids, reference ids = first
# appearance).
SAVECORE_CASES = [("g2000_L5_k2_lmd2", "g2000_L5", 2, 2), ("g2000_L5_kmax_lmd1", "g2000_L5", 0, 1),
                  ("c1_k3_lmd4", "c1", 3, 4), ("loader_k2_lmd1", "loader", 2, 1)]


def write_loader_dataset(ds: str) -> None:
    os.makedirs(ds, exist_ok=True)
    with open(os.path.join(ds, "mlg.conf"), "w") as f:
        for i in range(len(LOADER_LAYERS)):
            f.write(f"l{i}.txt\n")
    for i, text in enumerate(LOADER_LAYERS):
        with open(os.path.join(ds, f"l{i}.txt"), "w") as f:
            f.write(text)


def make_savecore_fixture(workdir: str) -> dict:
    """Runs the reference driver's kpe query per case (cwd run/, datasets in
    ../dataset/, the file lands in ../output/) and keeps its output file."""
    root = os.path.join(workdir, "savecore")
    for sub in ("dataset", "run", "output"):
        os.makedirs(os.path.join(root, sub), exist_ok=True)
    out = {}
    for case, dataset, k, lmd in SAVECORE_CASES:
        ds = os.path.join(root, "dataset", dataset)
        if not os.path.exists(ds):
            if dataset == "loader":
                write_loader_dataset(ds)
            else:
                MultilayerGraph.generate(GRAPHS[dataset]).write_text(ds)
        if k == 0:  # the largest k of this lmd, from the golden index
            k = int(np.load(os.path.join(HERE, dataset + ".npz"))["core_index"][lmd - 1].max())
        r = subprocess.run([REF_DRIVER, "-d", dataset, "-m", "kpe", "-k", str(k), "-lmd", str(lmd)],
                           cwd=os.path.join(root, "run"), capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"reference driver failed on {case}: {r.stderr[-2000:]}")
        path = os.path.join(root, "output", f"{dataset}_k_{k}_lmd_{lmd}_serial.txt")
        with open(path) as f:
            lines = f.read().splitlines()
        out[case + "_dataset"] = np.array(dataset)
        out[case + "_k"] = np.int64(k)
        out[case + "_lmd"] = np.int64(lmd)
        out[case + "_header"] = np.array(lines[0])
        out[case + "_labels"] = np.array([int(x) for x in lines[1:]], dtype=np.uint64)
    return out


def main() -> int:
    if "--savecore-only" in sys.argv:
        if not os.path.exists(REF_DRIVER):
            print("oracle/_ref/firmcore_ref is not built (make -C oracle ref)", file=sys.stderr)
            return 2
        with tempfile.TemporaryDirectory(prefix="hg_golden_") as work:
            fx = make_savecore_fixture(work)
        np.savez_compressed(os.path.join(HERE, "savecore.npz"), **fx)
        print("wrote savecore: " + ", ".join(f"{c}: {fx[c + '_header']} ({fx[c + '_labels'].size} labels)"
                                            for c, *_ in SAVECORE_CASES))
        return 0
    if not os.path.exists(HARNESS):
        print("oracle/_ref/ref_harness is not built (make -C oracle ref)", file=sys.stderr)
        return 2
    with tempfile.TemporaryDirectory(prefix="hg_golden_") as work:
        for name, spec in GRAPHS.items():
            fx = make_wdsci_fixture(name, spec, work)
            np.savez_compressed(os.path.join(HERE, f"{name}_wdsci.npz"), **fx)
            print(f"wrote {name}_wdsci: {fx['ci_1p0'].tolist()} printed {fx['ci_1p0_printed'].tolist()}")
        if "--wdsci-only" in sys.argv:
            return 0
        for name, spec in GRAPHS.items():
            fx = make_graph_fixture(name, spec, work)
            np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **fx)
            print(f"wrote {name}: n={fx['n']} L={fx['L']} cells={len(fx['cells'])} "
                  f"kmax={fx['core_index'].max(axis=1).tolist()} wds_1.0={fx['wds_1p0'][0]} "
                  f"{fx['wds_1p0_res'].tolist()}")
        fx = make_loader_fixture(work)
        np.savez_compressed(os.path.join(HERE, "loader.npz"), **fx)
        print(f"wrote loader: labels={fx['labels'].tolist()} index={fx['core_index'].tolist()}")
        fx = make_savecore_fixture(work)
        np.savez_compressed(os.path.join(HERE, "savecore.npz"), **fx)
    return 0


if __name__ == "__main__":
    sys.exit(main())
