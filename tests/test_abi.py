"""CPU tests of the C ABI boundary: both libraries load and export every
function include/*.h declares; argument checks that return before any
device call; the driver's flag handling."""
import ctypes as C
import os
import re
import subprocess

import pytest

from pargridcpu_amd import _native
from pargridcpu_amd import hipgrid as hg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = {"hipgrid.h": _native.ENGINE_LIB, "hg_host.h": _native.HOST_LIB}


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set(re.findall(r"\b(hgh?_[a-z_0-9]+)\s*\(", text))
    return sorted(n for n in names if not n.endswith("_t"))


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_exports_every_declared_symbol(header):
    lib = C.CDLL(HEADERS[header])
    names = declared_functions(header)
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{os.path.basename(HEADERS[header])} lacks {missing}"


def test_python_binding_covers_engine_abi():
    assert set(declared_functions("hipgrid.h")) == set(hg.SIGNATURES)


def test_create_argument_errors_without_device():
    lib = hg.lib()
    out = C.c_void_p()
    arr = (hg.LayerCSR * 1)()
    assert lib.hg_graph_create(None, 1, 10, 64, None, C.byref(out)) == -1
    assert lib.hg_graph_create(arr, 0, 10, 64, None, C.byref(out)) == -4
    assert lib.hg_graph_create(arr, 65, 10, 64, None, C.byref(out)) == -4
    assert lib.hg_graph_create(arr, 8, 1 << 29, 64, None, C.byref(out)) == -4  # n * L >= 2^32
    assert lib.hg_graph_create(arr, 1, (1 << 31) - 1, 64, None, C.byref(out)) == -4  # n >= 2^31 - 1
    assert b"2^31" in lib.hg_last_error()
    assert lib.hg_graph_create(arr, 1, 0, 64, None, C.byref(out)) == -1
    assert lib.hg_graph_create(arr, 1, 10, 16, None, C.byref(out)) == -1
    assert b"offset_bits" in lib.hg_last_error()
    assert lib.hg_decompose(None) == -1
    assert lib.hg_set_phase_timing(None, 1) == -1
    assert lib.hg_graph_destroy(None) == -1


def test_batch_argument_errors_without_device():
    lib = hg.lib()
    out = C.c_void_p()
    assert lib.hg_batch_create(None, 1, None, C.byref(out)) == -1
    descs = (hg.BatchGraph * 1)()
    assert lib.hg_batch_create(descs, 0, None, C.byref(out)) == -1
    assert b"count" in lib.hg_last_error()
    assert lib.hg_batch_decompose(None) == -1
    assert lib.hg_batch_destroy(None) == -1
    p = hg.BatchParams()
    assert lib.hg_batch_set_params(None, C.byref(p)) == -1


def test_driver_flags():
    exe = os.path.join(ROOT, "pargridcpu_amd", "bin", "hipgrid")
    assert os.path.exists(exe)
    r = subprocess.run([exe, "-d", "x", "-m", "ParGrid"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "unknown method" in r.stderr
    r = subprocess.run([exe, "-d", "missing", "-m", "HipGrid", "-root", "/nonexistent"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr
