"""Multilayer graph on the host: generator, reference text format, layer order.

Mirrors the reference's ``MultilayerGraph`` surface (Graph/MultilayerGraph.h:
10-52): ``LoadFromFile``, ``SetGraphOrder``, ``GetGraph(i)``, ``GetN``,
``getLayerNumber``, ``PrintStatistics``, ``LoadId2VtxMap``.  Storage is the
per-layer symmetric CSR of ``include/hg_host.h`` held by ``libhg_host.so``;
numpy views alias that memory without copies.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from ._native import load_host_lib

_lib = None


def _host():
    global _lib
    if _lib is None:
        _lib = load_host_lib()
    return _lib


class GenParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("n", C.c_int64),
        ("L", C.c_int32),
        ("gamma", C.c_double),
        ("rho", C.c_double),
        ("m_lo", C.c_int64),
        ("m_hi", C.c_int64),
        ("cap_deg", C.c_double),
    ]


@dataclass(frozen=True)
class GraphSpec:
    """Seeded Chung-Lu multilayer workload (see include/hg_host.h)."""

    name: str
    seed: int
    n: int
    L: int
    gamma: float
    rho: float
    m_lo: int
    m_hi: int
    cap_deg: float = 0.0
    order: int = 0

    def params(self) -> GenParams:
        return GenParams(self.seed, self.n, self.L, self.gamma, self.rho, self.m_lo, self.m_hi, self.cap_deg)


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = _host().hgh_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


class Layer:
    """One layer as CSR views (row_ptr int64[n+1], col uint32[m])."""

    def __init__(self, row_ptr: np.ndarray, col: np.ndarray):
        self.row_ptr = row_ptr
        self.col = col

    def degrees(self) -> np.ndarray:
        return np.diff(self.row_ptr)

    def getM(self) -> int:  # noqa: N802  (reference name, Graph.h:44)
        return int(self.col.shape[0])

    def GetAdjLst(self):  # noqa: N802  (reference name, Graph.h:53)
        """Row v as the reference's [deg, nbrs...] list (slow; for inspection)."""
        return [
            np.concatenate(([self.row_ptr[v + 1] - self.row_ptr[v]], self.col[self.row_ptr[v]:self.row_ptr[v + 1]]))
            for v in range(self.row_ptr.shape[0] - 1)
        ]


class MultilayerGraph:
    """Host multilayer graph; layer i of the ordered view is file layer order[i]."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        lib = _host()
        self.n = int(lib.hgh_graph_n(handle))
        self.L = int(lib.hgh_graph_layers(handle))
        self.order = list(range(self.L))
        self._layers = []
        for l in range(self.L):
            m = int(lib.hgh_graph_entries(handle, l))
            rp = np.ctypeslib.as_array(lib.hgh_graph_row_ptr(handle, l), shape=(self.n + 1,))
            if m > 0:
                col = np.ctypeslib.as_array(lib.hgh_graph_col(handle, l), shape=(m,))
            else:
                col = np.zeros(0, dtype=np.uint32)
            self._layers.append(Layer(rp, col))

    # -- construction ---------------------------------------------------------
    @classmethod
    def generate(cls, spec: GraphSpec) -> "MultilayerGraph":
        h = C.c_void_p()
        p = spec.params()
        _check(_host().hgh_generate(C.byref(p), C.byref(h)), "hgh_generate")
        g = cls(h)
        if spec.order:
            g.SetGraphOrder(spec.order)
        return g

    @classmethod
    def LoadFromFile(cls, input_path: str, write_vtx_map: bool = True) -> "MultilayerGraph":  # noqa: N802
        """Reference loader semantics (Graph/MultilayerGraph.cpp:65-166)."""
        h = C.c_void_p()
        _check(_host().hgh_load_text(os.fsencode(input_path), 1 if write_vtx_map else 0, C.byref(h)),
               "hgh_load_text")
        return cls(h)

    @classmethod
    def from_csr(cls, n: int, layers: list[tuple[np.ndarray, np.ndarray]]) -> "MultilayerGraph":
        """Wraps caller-provided symmetric CSR layers (kept by reference)."""
        obj = cls.__new__(cls)
        obj._h = None
        obj.n = int(n)
        obj.L = len(layers)
        obj.order = list(range(obj.L))
        obj._layers = [Layer(np.ascontiguousarray(rp, dtype=np.int64), np.ascontiguousarray(c, dtype=np.uint32))
                       for rp, c in layers]
        return obj

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.hgh_graph_free(h)
            self._h = None

    # -- reference surface ----------------------------------------------------
    def SetGraphOrder(self, o: int) -> None:  # noqa: N802  (MultilayerGraph.cpp:9-45)
        if o == 0:
            return
        if self._h is None:
            raise RuntimeError("SetGraphOrder needs a library-owned graph")
        order = (C.c_int32 * self.L)()
        _check(_host().hgh_layer_order(self._h, int(o), order), "hgh_layer_order")
        self.order = [int(x) for x in order]

    def GetGraph(self, i: int) -> Layer:  # noqa: N802  (MultilayerGraph.h:38-40)
        return self._layers[self.order[i]]

    def GetN(self) -> int:  # noqa: N802
        return self.n

    def getLayerNumber(self) -> int:  # noqa: N802
        return self.L

    def GetOrder(self) -> list[int]:  # noqa: N802
        return list(self.order)

    def PrintStatistics(self) -> str:  # noqa: N802  (MultilayerGraph.cpp:168-178)
        if self._h is None:
            m = [lay.getM() >> 1 for lay in self._layers]
            s = f"|L| = {self.L}, |V| = {self.n}, |E| = {sum(m) & 0xFFFFFFFF}, max_|E_i| = {max(m) if m else 0}"
        else:
            buf = C.create_string_buffer(256)
            _check(_host().hgh_statistics(self._h, buf, 256), "hgh_statistics")
            s = buf.value.decode()
        print(s)
        return s

    def LoadId2VtxMap(self) -> np.ndarray:  # noqa: N802  (MultilayerGraph.cpp:180-190)
        """Original label of every id (identity for generated graphs)."""
        if self._h is None:
            return np.arange(self.n, dtype=np.uint64)
        lib = _host()
        return np.fromiter((lib.hgh_graph_label(self._h, v) for v in range(self.n)), dtype=np.uint64, count=self.n)

    def write_text(self, directory: str) -> None:
        """Writes mlg.conf + one "u v" edge file per layer (file layer order)."""
        if self._h is None:
            raise RuntimeError("write_text needs a library-owned graph")
        os.makedirs(directory, exist_ok=True)
        _check(_host().hgh_write_text(self._h, os.fsencode(directory)), "hgh_write_text")

    # -- helpers ---------------------------------------------------------------
    def layers(self) -> list[Layer]:
        """Layers in the ordered view."""
        return [self.GetGraph(i) for i in range(self.L)]

    def directed_entries(self) -> int:
        return sum(lay.getM() for lay in self._layers)


# Benchmark workloads (seeded; SplitMix64 base seeds 0xC1.., Chung-Lu).
C1 = GraphSpec("c1-n20k-L7", 0xC1, 20_000, 7, 2.5, 0.5, 80_000, 80_000, 0.0, order=1)
C2 = GraphSpec("c2-n1M-L10", 0xC2, 1_000_000, 10, 2.2, 0.5, 300_000, 3_000_000, 0.0)
C3 = GraphSpec("c3-n100M-L8", 0xC3, 100_000_000, 8, 2.1, 0.3, 150_000_000, 150_000_000, 200_000.0)
WORKLOADS = {"c1": C1, "c2": C2, "c3": C3}

_MASK64 = (1 << 64) - 1


def _splitmix64(seed: int, t: int) -> int:
    """Output t of the SplitMix64 stream seeded with seed."""
    z = (seed + (t + 1) * 0x9E3779B97F4A7C15) & _MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
    return z ^ (z >> 31)


def _unit(seed: int, t: int) -> float:
    return (_splitmix64(seed, t) >> 11) * (1.0 / (1 << 53))


def batch_spec(g: int, base: int = 0xC4, n_lo: int = 1_000, n_hi: int = 200_000) -> GraphSpec:
    """Graph g of the batched workload: n log-uniform in [n_lo, n_hi], L
    uniform in {2..16}, average per-layer degree uniform in [2, 20], gamma
    uniform in [2.1, 3.0], 50% per-layer weight reshuffle; seed base<<32 | g
    draws the parameters (outputs 0..3 of its SplitMix64 stream) and seeds
    the generator."""
    import math

    seed = (base << 32) | g
    n = int(round(math.exp(math.log(n_lo) + _unit(seed, 0) * (math.log(n_hi) - math.log(n_lo)))))
    L = 2 + min(14, int(_unit(seed, 1) * 15))
    avg_deg = 2.0 + 18.0 * _unit(seed, 2)
    gamma = 2.1 + 0.9 * _unit(seed, 3)
    m = max(1, int(round(n * avg_deg / 2)))
    return GraphSpec(f"batch{base:x}-{g}", seed, n, L, gamma, 0.5, m, m, 0.0)


def batch_workload(count: int = 256, base: int = 0xC4) -> list[GraphSpec]:
    """The batched workload: `count` independent graphs (batch_spec)."""
    return [batch_spec(g, base) for g in range(count)]


def large_request_spec(r: int, base: int = 0xC5) -> GraphSpec:
    """The serving workload's large request: n = 5M, L = 12, 60M edges (5M
    sampled per layer), gamma 2.2, 50% reshuffle."""
    return GraphSpec(f"large{base:x}-{r}", (base << 32) | r, 5_000_000, 12, 2.2, 0.5, 5_000_000, 5_000_000, 0.0)
