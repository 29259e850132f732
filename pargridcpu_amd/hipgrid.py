"""HipGrid engine: ctypes binding of libhipgrid.so (include/hipgrid.h).

Host-side mirror of the reference's decomposition interface:

* ``FCSyncLeft.Execute(mg, tree)`` / ``ExecuteMix`` / ``FCSyncRight.Execute``
  / ``FCSyncMem.Execute`` (CoreParallel/FCSync*.h) all compute the same
  fcgrid; here ``HipGrid.Execute(mg, tree)`` fills an ``FCCoreTree`` with the
  flattened result (per-vertex core index for every lmd, cell lengths,
  k-shells) computed on the GPU.
* ``FCPathLevelLeft.Execute(mg, tree, wds=True, beta)`` followed by
  ``FCTree.WeightDenestSubgraph`` (Core/FCPathLevelLeft.cpp:311-377,
  Core/FCTree.cpp:69-82, main.cpp:402-424) is ``HipGrid.WeightDenestSubgraph``.

There is no CPU fallback: if libhipgrid.so or the GPU is missing, the calls
raise.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._native import ENGINE_LIB, NativeLibraryMissing, _open

HG_OK = 0
_ERRNAMES = {-1: "HG_ERR_ARG", -2: "HG_ERR_HIP", -3: "HG_ERR_NOMEM", -4: "HG_ERR_RANGE", -5: "HG_ERR_STATE",
             -6: "HG_ERR_DEVICE", -7: "HG_ERR_CANCELLED"}


class HgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class LayerCSR(C.Structure):
    _fields_ = [("row_ptr", C.c_void_p), ("col", C.c_void_p)]


class Wds(C.Structure):
    _fields_ = [("max_density", C.c_float), ("res_k", C.c_uint32), ("res_lmd", C.c_uint32), ("res_len", C.c_uint32),
                ("num_select_layer", C.c_int32), ("reserved", C.c_int32)]


class Timing(C.Structure):
    _fields_ = [("decompose_ms", C.c_double), ("init_ms", C.c_double), ("scan_ms", C.c_double),
                ("cascade_ms", C.c_double), ("densest_ms", C.c_double), ("levels", C.c_int64),
                ("launches", C.c_int64), ("binned", C.c_int64)]


class BatchGraph(C.Structure):
    _fields_ = [("layers", C.c_void_p), ("L", C.c_int32), ("offset_bits", C.c_int32), ("n", C.c_int64)]


class BatchParams(C.Structure):
    _fields_ = [("block", C.c_int32), ("unroll", C.c_int32), ("blocks_per_cu", C.c_int32), ("order", C.c_int32),
                ("use_graph", C.c_int32), ("reserved", C.c_int32)]

    def as_dict(self) -> dict:
        return {f: int(getattr(self, f)) for f, _ in self._fields_ if f != "reserved"}


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("large_entries", C.c_int64), ("lanes", C.c_int32), ("max_batch", C.c_int32),
                ("plain", C.c_int32), ("min_batch", C.c_int32), ("params", BatchParams), ("reuse_results", C.c_int32),
                ("device_queue", C.c_int32)]


class RequestInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("kind", C.c_int32), ("n", C.c_int64), ("L", C.c_int32),
                ("batch_size", C.c_int32), ("queue_ms", C.c_double), ("run_ms", C.c_double),
                ("latency_ms", C.c_double), ("wds", Wds)]


class ServerSummary(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64), ("batches", C.c_int64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
                ("span_s", C.c_double), ("peak_bytes", C.c_int64), ("cancelled", C.c_int64),
                ("resident_graphs", C.c_int64), ("resident_bytes", C.c_int64)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


REQ_DECOMPOSE = 0
REQ_DENSEST = 1

_lib = None

# Every symbol include/hipgrid.h declares, with its ctypes signature.
_P = C.c_void_p
SIGNATURES = {
    "hg_graph_create": ([C.POINTER(LayerCSR), C.c_int32, C.c_int64, C.c_int32, _P, C.POINTER(_P)], C.c_int),
    "hg_decompose": ([_P], C.c_int),
    "hg_core_index": ([_P, C.POINTER(C.c_int32)], C.c_int),
    "hg_core_index_device": ([_P, C.POINTER(_P)], C.c_int),
    "hg_cells": ([_P, C.POINTER(C.c_int64), C.c_int32, C.POINTER(C.c_int32)], C.c_int),
    "hg_shells": ([_P, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_uint32)], C.c_int),
    "hg_densest": ([_P, C.c_double, C.POINTER(Wds)], C.c_int),
    "hg_densest_coreindex": ([_P, C.c_double, C.POINTER(Wds), C.POINTER(Wds)], C.c_int),
    "hg_densest_noalloc": ([_P, C.c_double, C.POINTER(Wds)], C.c_int),
    "hg_cell_density": ([_P, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.c_int32], C.c_int),
    "hg_densest_vertices": ([_P, C.POINTER(C.c_uint32), C.c_int64, C.POINTER(C.c_int64)], C.c_int),
    "hg_last_timing": ([_P, C.POINTER(Timing)], C.c_int),
    "hg_set_phase_timing": ([_P, C.c_int32], C.c_int),
    "hg_graph_info": ([_P, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
                      C.c_int),
    "hg_graph_destroy": ([_P], C.c_int),
    "hg_last_error": ([], C.c_char_p),
    "hg_batch_create": ([C.POINTER(BatchGraph), C.c_int32, _P, C.POINTER(_P)], C.c_int),
    "hg_batch_set_params": ([_P, C.POINTER(BatchParams)], C.c_int),
    "hg_batch_get_params": ([_P, C.POINTER(BatchParams)], C.c_int),
    "hg_batch_tune": ([_P, C.c_int32, C.POINTER(BatchParams), C.POINTER(C.c_double)], C.c_int),
    "hg_batch_decompose": ([_P], C.c_int),
    "hg_batch_core_index": ([_P, C.c_int32, C.POINTER(C.c_int32)], C.c_int),
    "hg_batch_cells": ([_P, C.c_int32, C.POINTER(C.c_int64), C.c_int32, C.POINTER(C.c_int32)], C.c_int),
    "hg_batch_densest": ([_P, C.POINTER(C.c_int32), C.c_int32, C.c_double, C.POINTER(Wds)], C.c_int),
    "hg_batch_last_ms": ([_P, C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
    "hg_batch_info": ([_P, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
                      C.c_int),
    "hg_batch_destroy": ([_P], C.c_int),
    "hg_live_create": ([C.POINTER(BatchParams), C.c_int32, C.POINTER(_P)], C.c_int),
    "hg_batch_submit_live": ([_P, _P], C.c_int),
    "hg_batch_wait_live": ([_P, C.POINTER(C.c_int64)], C.c_int),
    "hg_live_stats": ([_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
    "hg_live_destroy": ([_P], C.c_int),
    "hg_server_create": ([C.POINTER(ServerConfig), C.POINTER(_P)], C.c_int),
    "hg_server_submit": ([_P, C.POINTER(BatchGraph), C.c_int32, C.c_double, C.POINTER(C.c_int64)], C.c_int),
    "hg_server_register": ([_P, C.POINTER(BatchGraph), C.POINTER(C.c_int64)], C.c_int),
    "hg_server_submit_resident": ([_P, C.c_int64, C.c_int32, C.c_double, C.POINTER(C.c_int64)], C.c_int),
    "hg_server_unregister": ([_P, C.c_int64], C.c_int),
    "hg_server_cancel": ([_P, C.c_int64], C.c_int),
    "hg_debug_fail_alloc": ([C.c_int64], C.c_int),
    "hg_server_wait": ([_P, C.c_int64, C.POINTER(RequestInfo)], C.c_int),
    "hg_server_result": ([_P, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
    "hg_server_release": ([_P, C.c_int64], C.c_int),
    "hg_server_drain": ([_P], C.c_int),
    "hg_server_stats": ([_P, C.POINTER(ServerSummary)], C.c_int),
    "hg_server_destroy": ([_P], C.c_int),
}


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        L = _open(ENGINE_LIB)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != HG_OK:
        raise HgError(rc, lib().hg_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Engine:
    """One graph resident on the GPU (hg_graph handle)."""

    def __init__(self, layers, n: int, stream: int | None = None, offset_bits: int = 64):
        """layers: sequence of objects with .row_ptr (int64/int32) and .col (uint32), in -o order."""
        self.n = int(n)
        self.L = len(layers)
        self._keep = []
        arr = (LayerCSR * max(self.L, 1))()
        want = np.int64 if offset_bits == 64 else np.int32
        for i, lay in enumerate(layers):
            rp = np.ascontiguousarray(lay.row_ptr, dtype=want)
            col = np.ascontiguousarray(lay.col, dtype=np.uint32)
            self._keep += [rp, col]
            arr[i].row_ptr = rp.ctypes.data
            arr[i].col = col.ctypes.data if col.size else rp.ctypes.data
        h = _P()
        _check(lib().hg_graph_create(arr, self.L, self.n, offset_bits, _P(stream) if stream else None, C.byref(h)))
        self._h = h
        self._keep = []  # the library copied the graph to the device

    @classmethod
    def from_graph(cls, mg, stream: int | None = None) -> "Engine":
        return cls(mg.layers(), mg.GetN(), stream=stream)

    def close(self) -> None:
        if getattr(self, "_h", None):
            _check(lib().hg_graph_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- calls ----------------------------------------------------------------
    def decompose(self) -> None:
        _check(lib().hg_decompose(self._h))

    def core_index(self) -> np.ndarray:
        out = np.empty((self.L, self.n), dtype=np.int32)
        _check(lib().hg_core_index(self._h, _ptr(out, C.c_int32)))
        return out

    def core_index_device_ptr(self) -> int:
        p = _P()
        _check(lib().hg_core_index_device(self._h, C.byref(p)))
        return int(p.value or 0)

    def cells(self, cap: int | None = None):
        kmax = np.zeros(self.L, dtype=np.int32)
        _check(lib().hg_cells(self._h, None, 0, _ptr(kmax, C.c_int32)))
        if cap is None:
            cap = int(kmax.max()) + 2
        cl = np.zeros((self.L, cap), dtype=np.int64)
        _check(lib().hg_cells(self._h, _ptr(cl, C.c_int64), cap, None))
        return cl, kmax

    def shells(self):
        _, kmax = self.cells()
        kcap = int(kmax.max()) + 1
        off = np.zeros(self.L * (kcap + 1) + 1, dtype=np.int64)
        vtx = np.zeros(self.n * self.L, dtype=np.uint32)
        _check(lib().hg_shells(self._h, kcap, _ptr(off, C.c_int64), _ptr(vtx, C.c_uint32)))
        return off, vtx, kcap

    def densest(self, beta: float = 1.0) -> Wds:
        w = Wds()
        _check(lib().hg_densest(self._h, float(beta), C.byref(w)))
        return w

    def densest_coreindex(self, beta: float = 1.0):
        """-m WdsCoreIndex: (printed answer, best cell per lmd)."""
        ans = Wds()
        per = (Wds * self.L)()
        _check(lib().hg_densest_coreindex(self._h, float(beta), C.byref(ans), per))
        return ans, [per[i] for i in range(self.L)]

    def cell_density(self, cap: int):
        d = np.zeros((self.L, cap), dtype=np.float64)
        s = np.zeros((self.L, cap), dtype=np.int32)
        _check(lib().hg_cell_density(self._h, _ptr(d, C.c_double), _ptr(s, C.c_int32), cap))
        return d, s

    def densest_vertices(self) -> np.ndarray:
        cnt = C.c_int64(0)
        _check(lib().hg_densest_vertices(self._h, None, 0, C.byref(cnt)))
        out = np.zeros(max(cnt.value, 1), dtype=np.uint32)
        _check(lib().hg_densest_vertices(self._h, _ptr(out, C.c_uint32), out.size, C.byref(cnt)))
        return out[:cnt.value]

    def timing(self) -> Timing:
        t = Timing()
        _check(lib().hg_last_timing(self._h, C.byref(t)))
        return t

    def set_phase_timing(self, on: bool = True) -> None:
        """Events around every scan and cascade of later decompositions, so
        timing().scan_ms / cascade_ms are measured (else -1)."""
        _check(lib().hg_set_phase_timing(self._h, 1 if on else 0))

    def info(self) -> dict:
        n, L, D, b = C.c_int64(), C.c_int32(), C.c_int64(), C.c_int64()
        _check(lib().hg_graph_info(self._h, C.byref(n), C.byref(L), C.byref(D), C.byref(b)))
        return {"n": n.value, "L": L.value, "directed_entries": D.value, "device_bytes": b.value}


@dataclass
class FCCoreTree:
    """Flattened fcgrid: what the reference's coreNodeP tree records per cell.

    core_index[lmd-1][v]; cell (k, lmd) is non-empty iff 1 <= k <= kmax[lmd-1];
    its length (coreNodeP::length) is cell_len[lmd-1][k]."""

    core_index: np.ndarray | None = None
    cell_len: np.ndarray | None = None
    kmax: np.ndarray | None = None

    def length(self, k: int, lmd: int) -> int:
        if lmd < 1 or lmd > self.kmax.shape[0] or k < 1 or k > self.kmax[lmd - 1]:
            return 0
        return int(self.cell_len[lmd - 1][k])

    def getCoreByKAndLmd(self, k: int, lmd: int) -> np.ndarray | None:  # noqa: N802
        """Vertex ids of S(k, lmd), ascending; None for an empty cell
        (getCoreByKAndLmdByLeft/Right return nullptr, Core/FCTree.cpp:106-136)."""
        if self.length(k, lmd) == 0:
            return None
        return np.nonzero(self.core_index[lmd - 1] >= k)[0].astype(np.uint32)

    def count(self) -> int:
        """Number of non-empty cells (the reference's printed `count`)."""
        return int(self.kmax.sum())


class HipGrid:
    """-m HipGrid: drop-in for -m ParGrid / ParGridPlus / ParGridright / CoreMem."""

    @staticmethod
    def Execute(mg, tree: FCCoreTree | None = None, stream: int | None = None) -> FCCoreTree:  # noqa: N802
        tree = tree if tree is not None else FCCoreTree()
        with Engine.from_graph(mg, stream=stream) as eng:
            eng.decompose()
            tree.core_index = eng.core_index()
            tree.cell_len, tree.kmax = eng.cells()
        return tree

    ExecuteMix = Execute

    @staticmethod
    def WeightDenestSubgraph(mg, beta: float = 1.0, stream: int | None = None):  # noqa: N802
        """-m WdsList result: (maximum_density float32, res_k, res_lmd, res_len, vertex ids)."""
        with Engine.from_graph(mg, stream=stream) as eng:
            eng.decompose()
            w = eng.densest(beta)
            verts = eng.densest_vertices()
        return np.float32(w.max_density), int(w.res_k), int(w.res_lmd), int(w.res_len), verts


def _graph_desc(g, offset_bits: int = 64):
    """(BatchGraph, arrays to keep alive) for an object with .layers() and .n."""
    want = np.int64 if offset_bits == 64 else np.int32
    layers, n = (g.layers(), g.n) if hasattr(g, "layers") else g
    keep = []
    arr = (LayerCSR * max(len(layers), 1))()
    for l, lay in enumerate(layers):
        rp = np.ascontiguousarray(lay.row_ptr, dtype=want)
        col = np.ascontiguousarray(lay.col, dtype=np.uint32)
        keep += [rp, col]
        arr[l].row_ptr = rp.ctypes.data
        arr[l].col = col.ctypes.data if col.size else rp.ctypes.data
    keep.append(arr)
    d = BatchGraph()
    d.layers = C.cast(arr, C.c_void_p)
    d.L = len(layers)
    d.offset_bits = offset_bits
    d.n = int(n)
    return d, keep


class Batch:
    """Many graphs decomposed in one persistent launch (hg_batch handle).

    graphs: objects with .layers() (CSR layers in -o order) and .n, or
    (layers, n) pairs."""

    def __init__(self, graphs, stream: int | None = None, offset_bits: int = 64):
        keep = []
        descs = (BatchGraph * max(len(graphs), 1))()
        self.shapes = []
        for i, g in enumerate(graphs):
            d, k = _graph_desc(g, offset_bits)
            keep.append(k)
            descs[i] = d
            self.shapes.append((int(d.n), int(d.L)))
        h = _P()
        _check(lib().hg_batch_create(descs, len(graphs), _P(stream) if stream else None, C.byref(h)))
        self._h = h
        self.count = len(graphs)

    def close(self) -> None:
        if getattr(self, "_h", None):
            _check(lib().hg_batch_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_params(self, **kw) -> None:
        p = self.params()
        for k, v in kw.items():
            setattr(p, k, int(v))
        _check(lib().hg_batch_set_params(self._h, C.byref(p)))

    def params(self) -> BatchParams:
        p = BatchParams()
        _check(lib().hg_batch_get_params(self._h, C.byref(p)))
        return p

    def tune(self, reps: int = 3):
        p, ms = BatchParams(), C.c_double()
        _check(lib().hg_batch_tune(self._h, int(reps), C.byref(p), C.byref(ms)))
        return p.as_dict(), ms.value

    def decompose(self) -> float:
        """Runs the launch; returns its device time in ms."""
        _check(lib().hg_batch_decompose(self._h))
        return self.last_ms()[0]

    def submit_live(self, q: "Live") -> None:
        """Appends every item of the batch to the live queue q (returns at once)."""
        _check(lib().hg_batch_submit_live(self._h, q._h))

    def wait_live(self) -> int:
        """Waits for the live submission; returns how many of its items were
        taken by launches already on the GPU when it was submitted."""
        t = C.c_int64()
        _check(lib().hg_batch_wait_live(self._h, C.byref(t)))
        return t.value

    def last_ms(self):
        a, b = C.c_double(), C.c_double()
        _check(lib().hg_batch_last_ms(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def core_index(self, g: int) -> np.ndarray:
        n, L = self.shapes[g]
        out = np.empty((L, n), dtype=np.int32)
        _check(lib().hg_batch_core_index(self._h, int(g), _ptr(out, C.c_int32)))
        return out

    def cells(self, g: int, cap: int | None = None):
        n, L = self.shapes[g]
        kmax = np.zeros(L, dtype=np.int32)
        _check(lib().hg_batch_cells(self._h, int(g), None, 0, _ptr(kmax, C.c_int32)))
        if cap is None:
            cap = int(kmax.max()) + 2
        cl = np.zeros((L, cap), dtype=np.int64)
        _check(lib().hg_batch_cells(self._h, int(g), _ptr(cl, C.c_int64), cap, None))
        return cl, kmax

    def densest(self, graphs, beta: float = 1.0) -> list:
        ids = np.ascontiguousarray(list(graphs), dtype=np.int32)
        out = (Wds * max(len(ids), 1))()
        _check(lib().hg_batch_densest(self._h, _ptr(ids, C.c_int32), len(ids), float(beta), out))
        return [out[i] for i in range(len(ids))]

    def info(self) -> dict:
        c, it, D, b = C.c_int32(), C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().hg_batch_info(self._h, C.byref(c), C.byref(it), C.byref(D), C.byref(b)))
        return {"graphs": c.value, "items": it.value, "directed_entries": D.value, "device_bytes": b.value}


class Live:
    """Device work queue shared by batches (hg_live handle): a batch
    submitted while others run joins the workgroups already on the GPU."""

    def __init__(self, block: int = 1024, unroll: int = 4, blocks_per_cu: int = 1, slots: int = 0):
        p = BatchParams()
        p.block, p.unroll, p.blocks_per_cu, p.order, p.use_graph = block, unroll, blocks_per_cu, 0, 0
        h = _P()
        _check(lib().hg_live_create(C.byref(p), int(slots), C.byref(h)))
        self._h = h

    def stats(self) -> dict:
        d, r, n = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().hg_live_stats(self._h, C.byref(d), C.byref(r), C.byref(n)))
        return {"items_done": d.value, "items_taken_running": r.value, "launches": n.value}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _check(lib().hg_live_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """Request loop over one GPU (hg_server handle): submit graphs for
    decomposition or densest-subgraph search, wait for their answers."""

    def __init__(self, hbm_budget: int = 200 << 30, lanes: int = 4, max_batch: int = 64,
                 large_entries: int = 4_000_000, plain: bool = False, params: dict | None = None,
                 min_batch: int = 0, reuse_results: bool = True, device_queue: bool = False):
        cfg = ServerConfig()
        cfg.hbm_budget = int(hbm_budget)
        cfg.large_entries = int(large_entries)
        cfg.lanes = int(lanes)
        cfg.max_batch = int(max_batch)
        cfg.plain = 1 if plain else 0
        cfg.min_batch = int(min_batch)
        cfg.reuse_results = 1 if reuse_results else 0
        cfg.device_queue = 1 if device_queue else 0
        for k, v in (params or {}).items():
            setattr(cfg.params, k, int(v))
        h = _P()
        _check(lib().hg_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._shapes = {}

    def submit(self, g, kind: int = REQ_DECOMPOSE, beta: float = 1.0) -> int:
        d, keep = _graph_desc(g)
        rid = C.c_int64()
        _check(lib().hg_server_submit(self._h, C.byref(d), int(kind), float(beta), C.byref(rid)))
        self._shapes[rid.value] = (int(d.n), int(d.L))
        return rid.value

    def register(self, g) -> int:
        """Uploads g once and keeps it resident; returns its graph id."""
        d, keep = _graph_desc(g)
        gid = C.c_int64()
        _check(lib().hg_server_register(self._h, C.byref(d), C.byref(gid)))
        self._gshapes = getattr(self, "_gshapes", {})
        self._gshapes[gid.value] = (int(d.n), int(d.L))
        return gid.value

    def submit_resident(self, gid: int, kind: int = REQ_DECOMPOSE, beta: float = 1.0) -> int:
        rid = C.c_int64()
        _check(lib().hg_server_submit_resident(self._h, int(gid), int(kind), float(beta), C.byref(rid)))
        self._shapes[rid.value] = self._gshapes[gid]
        return rid.value

    def unregister(self, gid: int) -> None:
        _check(lib().hg_server_unregister(self._h, int(gid)))

    def cancel(self, rid: int) -> None:
        _check(lib().hg_server_cancel(self._h, int(rid)))

    def wait(self, rid: int) -> RequestInfo:
        info = RequestInfo()
        _check(lib().hg_server_wait(self._h, int(rid), C.byref(info)))
        return info

    def result(self, rid: int):
        n, L = self._shapes[rid]
        ci = np.empty((L, n), dtype=np.int32)
        kmax = np.empty(L, dtype=np.int32)
        _check(lib().hg_server_result(self._h, int(rid), _ptr(ci, C.c_int32), _ptr(kmax, C.c_int32)))
        return ci, kmax

    def release(self, rid: int) -> None:
        _check(lib().hg_server_release(self._h, int(rid)))
        self._shapes.pop(rid, None)

    def drain(self) -> None:
        _check(lib().hg_server_drain(self._h))

    def stats(self) -> dict:
        s = ServerSummary()
        _check(lib().hg_server_stats(self._h, C.byref(s)))
        return s.as_dict()

    def close(self) -> None:
        if getattr(self, "_h", None):
            _check(lib().hg_server_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


__all__ = ["Batch", "Engine", "Live", "Server", "FCCoreTree", "HipGrid", "HgError", "NativeLibraryMissing", "SIGNATURES"]
