"""ctypes binding of oracle/liboracle.so (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  See firmcore_oracle.c for what each
function restates and how it is pinned.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_DIR, "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built; run __graft_entry__.build()")
        _lib = C.CDLL(_LIB_PATH)
        _lib.orc_check_lmd.restype = C.c_int64
    return _lib


def _layer_ptrs(layers):
    rps = [np.ascontiguousarray(l.row_ptr, dtype=np.int64) for l in layers]
    cols = [np.ascontiguousarray(l.col, dtype=np.uint32) for l in layers]
    RP = (C.POINTER(C.c_int64) * len(rps))(*[r.ctypes.data_as(C.POINTER(C.c_int64)) for r in rps])
    CL = (C.POINTER(C.c_uint32) * len(cols))(*[c.ctypes.data_as(C.POINTER(C.c_uint32)) for c in cols])
    return RP, CL, (rps, cols)


def _i32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _i64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def pargrid(n, layers, cap=None):
    """ParGrid restatement -> (core_index[L][n], cell_len[L][cap], kmax[L])."""
    L = len(layers)
    RP, CL, keep = _layer_ptrs(layers)
    if cap is None:
        cap = int(max((l.degrees().max() if l.row_ptr.size > 1 else 0) for l in layers)) + 2
    ci = np.zeros((L, n), dtype=np.int32)
    cells = np.zeros((L, cap), dtype=np.int64)
    kmax = np.zeros(L, dtype=np.int32)
    rc = lib().orc_pargrid(C.c_int64(n), C.c_int32(L), RP, CL, _i32p(ci), _i64p(cells), C.c_int32(cap), _i32p(kmax))
    if rc != 0:
        raise RuntimeError(f"orc_pargrid failed {rc}")
    return ci, cells, kmax


def tree(n, layers, right_spine=False, cap=None):
    """lpe (right_spine=False) / kpe restatement -> dict."""
    L = len(layers)
    RP, CL, keep = _layer_ptrs(layers)
    if cap is None:
        cap = int(max((l.degrees().max() if l.row_ptr.size > 1 else 0) for l in layers)) + 2
    ci = np.zeros((L, n), dtype=np.int32)
    cells = np.zeros((L, cap), dtype=np.int64)
    kmax = np.zeros(L, dtype=np.int32)
    count = C.c_int64(0)
    sfc = C.c_int64(0)
    efc = np.zeros(cap, dtype=np.int64)
    rc = lib().orc_tree(C.c_int64(n), C.c_int32(L), RP, CL, C.c_int32(1 if right_spine else 0), _i32p(ci),
                        _i64p(cells), C.c_int32(cap), _i32p(kmax), C.byref(count), C.byref(sfc), _i64p(efc),
                        C.c_int32(cap))
    if rc != 0:
        raise RuntimeError(f"orc_tree failed {rc}")
    return {"core_index": ci, "cell_len": cells, "kmax": kmax, "count": count.value,
            "sum_num_of_edge_fc": sfc.value, "edge_fc_k1": efc}


def coreindex(n, layers):
    L = len(layers)
    RP, CL, keep = _layer_ptrs(layers)
    ci = np.zeros((L, n), dtype=np.int32)
    rc = lib().orc_coreindex(C.c_int64(n), C.c_int32(L), RP, CL, _i32p(ci))
    if rc != 0:
        raise RuntimeError(f"orc_coreindex failed {rc}")
    return ci


def coreindex_lmd(n, layers, lmd, out=None):
    """One lmd of the CoreIndex restatement (one iteration of ParCoreIndex's
    parallel loop, CoreIndex.cpp:217-300) into out (int32[n])."""
    L = len(layers)
    RP, CL, keep = _layer_ptrs(layers)
    if out is None:
        out = np.zeros(n, dtype=np.int32)
    rc = lib().orc_coreindex_lmd(C.c_int64(n), C.c_int32(L), RP, CL, C.c_int32(lmd), _i32p(out))
    if rc != 0:
        raise RuntimeError(f"orc_coreindex_lmd failed {rc}")
    return out


def merged(n, layers):
    """Vertex-major merged adjacency (orc_merge): off int64[n+1], adj
    uint32[D] with entries u * L + l."""
    L = len(layers)
    RP, CL, keep = _layer_ptrs(layers)
    D = int(sum(int(l.row_ptr[-1] - l.row_ptr[0]) for l in layers))
    off = np.empty(n + 1, dtype=np.int64)
    adj = np.empty(max(D, 1), dtype=np.uint32)
    rc = lib().orc_merge(C.c_int64(n), C.c_int32(L), RP, CL, _i64p(off), adj.ctypes.data_as(C.POINTER(C.c_uint32)))
    if rc != 0:
        raise RuntimeError(f"orc_merge failed {rc}")
    return off, adj


def coreindex_parallel(n, layers, threads=None, lmds=None, progress=None):
    """CoreIndex restatement with the lmds run concurrently, one host thread
    each (as ParCoreIndex runs them, CoreIndex.cpp:252) over one shared
    merged adjacency; ctypes releases the GIL during the calls.  Returns
    core_index[L][n] (rows not in `lmds` stay 0).  progress(lmd, seconds)
    is called as each lmd finishes."""
    import concurrent.futures as cf
    import time

    L = len(layers)
    off, adj = merged(n, layers)
    ci = np.zeros((L, n), dtype=np.int32)
    todo = list(range(1, L + 1)) if lmds is None else list(lmds)
    threads = threads or min(len(todo), os.cpu_count() or 1, 16)
    t0 = time.perf_counter()

    def one(lmd):
        rc = lib().orc_coreindex_merged_lmd(C.c_int64(n), C.c_int32(L), _i64p(off),
                                            adj.ctypes.data_as(C.POINTER(C.c_uint32)), C.c_int32(lmd),
                                            _i32p(ci[lmd - 1]))
        if rc != 0:
            raise RuntimeError(f"orc_coreindex_merged_lmd(lmd={lmd}) failed {rc}")
        if progress:
            progress(lmd, time.perf_counter() - t0)

    with cf.ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        list(ex.map(one, todo))
    return ci


def check_local(n, layers, core_index, threads=None, lmds=None):
    """Local FirmCore conditions (orc_check_lmd) on every row of core_index;
    returns {lmd: (violations, first violating id)} for the rows checked."""
    import concurrent.futures as cf

    L = len(layers)
    RP, CL, keep = _layer_ptrs(layers)
    ci = np.ascontiguousarray(core_index, dtype=np.int32)
    todo = list(range(1, L + 1)) if lmds is None else list(lmds)
    threads = threads or min(16, os.cpu_count() or 1)
    parts = max(1, threads)
    bounds = [n * i // parts for i in range(parts + 1)]
    jobs = [(lmd, bounds[i], bounds[i + 1]) for lmd in todo for i in range(parts) if bounds[i + 1] > bounds[i]]

    def one(job):
        lmd, v0, v1 = job
        fb = C.c_int64(-1)
        bad = lib().orc_check_lmd(C.c_int64(n), C.c_int32(L), RP, CL, C.c_int32(lmd), _i32p(ci[lmd - 1]),
                                  C.c_int64(v0), C.c_int64(v1), C.byref(fb))
        if bad < 0:
            raise RuntimeError("orc_check_lmd: bad arguments")
        return lmd, bad, fb.value

    out = {lmd: (0, -1) for lmd in todo}
    with cf.ThreadPoolExecutor(max_workers=parts) as ex:
        for lmd, bad, fb in ex.map(one, jobs):
            b0, f0 = out[lmd]
            out[lmd] = (b0 + bad, f0 if f0 >= 0 else fb)
    return out


def bruteforce(n, layers):
    L = len(layers)
    RP, CL, keep = _layer_ptrs(layers)
    ci = np.zeros((L, n), dtype=np.int32)
    rc = lib().orc_bruteforce(C.c_int64(n), C.c_int32(L), RP, CL, _i32p(ci))
    if rc != 0:
        raise RuntimeError(f"orc_bruteforce failed {rc}")
    return ci


def pow_table(L, beta):
    """pow(l+1, beta) for l = 0..L-1 via the C library pow, as the reference."""
    return np.array([math.pow(float(l + 1), float(beta)) for l in range(L)], dtype=np.float64)


def wds(n, layers, beta=1.0, cap=None):
    """WdsList restatement -> dict(max_density float32, res_k, res_lmd, res_len, cell_density, cell_sel)."""
    L = len(layers)
    RP, CL, keep = _layer_ptrs(layers)
    if cap is None:
        cap = int(max((l.degrees().max() if l.row_ptr.size > 1 else 0) for l in layers)) + 2
    pw = pow_table(L, beta)
    md = C.c_float(0)
    rk, rl, rn = C.c_uint32(0), C.c_uint32(0), C.c_uint32(0)
    dens = np.zeros((L, cap), dtype=np.float64)
    sel = np.zeros((L, cap), dtype=np.int32)
    rc = lib().orc_wds(C.c_int64(n), C.c_int32(L), RP, CL, pw.ctypes.data_as(C.POINTER(C.c_double)), C.byref(md),
                       C.byref(rk), C.byref(rl), C.byref(rn), dens.ctypes.data_as(C.POINTER(C.c_double)), _i32p(sel),
                       C.c_int32(cap))
    if rc != 0:
        raise RuntimeError(f"orc_wds failed {rc}")
    return {"max_density": np.float32(md.value), "res_k": rk.value, "res_lmd": rl.value, "res_len": rn.value,
            "cell_density": dens, "cell_sel": sel}


def wds_coreindex(n, layers, beta=1.0):
    """WdsCoreIndex restatement: per-lmd (density float32, res_k, res_lmd) as
    get_wds leaves them, and the answer CoreIndex::WdsCoreIndex prints."""
    L = len(layers)
    RP, CL, keep = _layer_ptrs(layers)
    dens = np.zeros(L, dtype=np.float32)
    rk = np.zeros(L, dtype=np.uint32)
    rl = np.zeros(L, dtype=np.uint32)
    rc = lib().orc_wds_coreindex(C.c_int64(n), C.c_int32(L), RP, CL, C.c_float(beta),
                                 dens.ctypes.data_as(C.POINTER(C.c_float)), rk.ctypes.data_as(C.POINTER(C.c_uint32)),
                                 rl.ctypes.data_as(C.POINTER(C.c_uint32)))
    if rc != 0:
        raise RuntimeError(f"orc_wds_coreindex failed {rc}")
    return {"density": dens, "res_k": rk, "res_lmd": rl, "answer": wds_coreindex_answer(dens, rk, rl)}


def wds_coreindex_answer(dens, rk, rl):
    """CoreIndex::WdsCoreIndex's printed answer (CoreIndexParallel/CoreIndex.cpp:
    448-478): records [0..L] with record lmd = get_wds's result for lmd and
    record 0 never written; std::sort of records [0, L) by density,
    descending, prints record 0 after the sort.  So lmd = L takes no part,
    and with L <= 16 (insertion sort, stable) the first record with the
    largest density wins.  Record 0 is read uninitialised by the reference;
    it is taken as zeros here (for L = 1 the reference prints garbage)."""
    recs = [(np.float32(0.0), 0, 0)] + [(np.float32(dens[i]), int(rk[i]), int(rl[i])) for i in range(len(dens))]
    recs = recs[:len(dens)]  # records [0, L)
    best = recs[0]
    for r in recs[1:]:
        if r[0] > best[0]:
            best = r
    return {"max_density": best[0], "res_k": best[1], "res_lmd": best[2]}
