"""Small hand-built multilayer graphs for closed-form and edge-case tests."""
from __future__ import annotations

import numpy as np

from pargridcpu_amd.graph import MultilayerGraph


def csr_from_edges(n: int, edges) -> tuple[np.ndarray, np.ndarray]:
    """Symmetric, sorted, de-duplicated CSR without self-loops."""
    e = np.asarray(list(edges), dtype=np.int64).reshape(-1, 2)
    if e.size:
        e = e[e[:, 0] != e[:, 1]]
        e = np.concatenate([e, e[:, ::-1]])
        e = np.unique(e, axis=0)
    deg = np.bincount(e[:, 0], minlength=n) if e.size else np.zeros(n, dtype=np.int64)
    rp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(deg, out=rp[1:])
    col = e[:, 1].astype(np.uint32) if e.size else np.zeros(0, dtype=np.uint32)
    return rp, col


def graph(n: int, layer_edges) -> MultilayerGraph:
    return MultilayerGraph.from_csr(n, [csr_from_edges(n, es) for es in layer_edges])


def clique(vs):
    vs = list(vs)
    return [(a, b) for i, a in enumerate(vs) for b in vs[i + 1:]]


def star(center, leaves):
    return [(center, x) for x in leaves]


def cycle(vs):
    vs = list(vs)
    return [(vs[i], vs[(i + 1) % len(vs)]) for i in range(len(vs))]


def path(vs):
    vs = list(vs)
    return [(vs[i], vs[i + 1]) for i in range(len(vs) - 1)]


def kcore_numbers(n: int, edges) -> np.ndarray:
    """Classic k-core numbers by repeated minimum-degree removal (Matula-Beck);
    vertices without edges get 0."""
    rp, col = csr_from_edges(n, edges)
    deg = np.diff(rp).astype(np.int64)
    alive = np.ones(n, dtype=bool)
    core = np.zeros(n, dtype=np.int64)
    k = 0
    for _ in range(n):
        cand = np.where(alive)[0]
        v = cand[np.argmin(deg[cand])]
        k = max(k, int(deg[v]))
        core[v] = k
        alive[v] = False
        for u in col[rp[v]:rp[v + 1]]:
            if alive[u]:
                deg[u] -= 1
    return core


def random_layers(n: int, L: int, m: int, seed: int):
    rng = np.random.default_rng(seed)
    return [list(map(tuple, rng.integers(0, n, size=(m, 2)))) for _ in range(L)]
