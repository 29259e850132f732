"""Locates and binds the in-tree native libraries (built by __graft_entry__.build)."""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(PKG_DIR, "libhg_host.so")
ENGINE_LIB = os.path.join(PKG_DIR, "libhipgrid.so")


class NativeLibraryMissing(RuntimeError):
    pass


def _open(path: str) -> C.CDLL:
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{os.path.basename(path)} is not built (expected at {path}); run __graft_entry__.build()")
    return C.CDLL(path)


def load_host_lib() -> C.CDLL:
    lib = _open(HOST_LIB)
    P = C.c_void_p
    lib.hgh_generate.argtypes = [P, C.POINTER(P)]
    lib.hgh_generate.restype = C.c_int
    lib.hgh_load_text.argtypes = [C.c_char_p, C.c_int32, C.POINTER(P)]
    lib.hgh_load_text.restype = C.c_int
    lib.hgh_write_text.argtypes = [P, C.c_char_p]
    lib.hgh_write_text.restype = C.c_int
    lib.hgh_layer_order.argtypes = [P, C.c_int32, C.POINTER(C.c_int32)]
    lib.hgh_layer_order.restype = C.c_int
    lib.hgh_permute_layers.argtypes = [P, C.POINTER(C.c_int32), C.POINTER(P)]
    lib.hgh_permute_layers.restype = C.c_int
    lib.hgh_graph_n.argtypes = [P]
    lib.hgh_graph_n.restype = C.c_int64
    lib.hgh_graph_layers.argtypes = [P]
    lib.hgh_graph_layers.restype = C.c_int32
    lib.hgh_graph_row_ptr.argtypes = [P, C.c_int32]
    lib.hgh_graph_row_ptr.restype = C.POINTER(C.c_int64)
    lib.hgh_graph_col.argtypes = [P, C.c_int32]
    lib.hgh_graph_col.restype = C.POINTER(C.c_uint32)
    lib.hgh_graph_entries.argtypes = [P, C.c_int32]
    lib.hgh_graph_entries.restype = C.c_int64
    lib.hgh_graph_label.argtypes = [P, C.c_int64]
    lib.hgh_graph_label.restype = C.c_uint64
    lib.hgh_graph_free.argtypes = [P]
    lib.hgh_graph_free.restype = None
    lib.hgh_statistics.argtypes = [P, C.c_char_p, C.c_int32]
    lib.hgh_statistics.restype = C.c_int
    lib.hgh_peak_rss_mb.argtypes = []
    lib.hgh_peak_rss_mb.restype = C.c_double
    lib.hgh_last_error.argtypes = []
    lib.hgh_last_error.restype = C.c_char_p
    return lib
