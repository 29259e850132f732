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
             -6: "HG_ERR_DEVICE"}


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
                ("launches", C.c_int64)]


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
    "hg_cell_density": ([_P, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.c_int32], C.c_int),
    "hg_densest_vertices": ([_P, C.POINTER(C.c_uint32), C.c_int64, C.POINTER(C.c_int64)], C.c_int),
    "hg_last_timing": ([_P, C.POINTER(Timing)], C.c_int),
    "hg_graph_info": ([_P, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
                      C.c_int),
    "hg_graph_destroy": ([_P], C.c_int),
    "hg_last_error": ([], C.c_char_p),
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


__all__ = ["Engine", "FCCoreTree", "HipGrid", "HgError", "NativeLibraryMissing", "SIGNATURES"]
