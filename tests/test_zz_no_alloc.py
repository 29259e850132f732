"""GPU test of hg_densest_noalloc (the densest-subgraph pass with its buffers
in a fixed device array): through the C ABI, the same answer, per-cell
densities and vertex set as hg_densest, on two calls in a row with the
second on a smaller graph, then again on the first; a graph whose pass does
not fit the array is refused with HG_ERR_RANGE."""
import ctypes as C
import dataclasses
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pargridcpu_amd.graph import C1, MultilayerGraph
from pargridcpu_amd.hipgrid import Engine, Wds, lib
from tests import graphs as G

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 256 * 1024  # HG_NOALLOC_SCRATCH_BYTES in include/hipgrid.h


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def spec(n, L, m, seed):
    return dataclasses.replace(C1, name=f"na{n}", seed=seed, n=n, L=L, m_lo=m, m_hi=m, order=0)


def answer(w):
    return (np.float32(w.max_density), w.res_k, w.res_lmd, w.res_len, w.num_select_layer)


def both(eng, beta):
    """(hg_densest answer, its cell densities, its vertex set) and the same
    from hg_densest_noalloc, both through the C ABI."""
    ref = Wds()
    assert lib().hg_densest(eng._h, C.c_double(beta), C.byref(ref)) == 0
    cap = int(eng.cells()[1].max()) + 2
    d_ref, s_ref = eng.cell_density(cap)
    v_ref = eng.densest_vertices()
    got = Wds()
    rc = lib().hg_densest_noalloc(eng._h, C.c_double(beta), C.byref(got))
    assert rc == 0, lib().hg_last_error()
    d_got, s_got = eng.cell_density(cap)
    v_got = eng.densest_vertices()
    return (answer(ref), d_ref, s_ref, v_ref), (answer(got), d_got, s_got, v_got)


def test_noalloc_matches_densest_on_consecutive_calls():
    big = MultilayerGraph.generate(spec(20000, 7, 80000, 0xA1))
    small = MultilayerGraph.generate(spec(3000, 4, 9000, 0xA2))
    with Engine.from_graph(big) as e1, Engine.from_graph(small) as e2:
        e1.decompose()
        e2.decompose()
        for step, (eng, mg, beta) in enumerate([(e1, big, 1.0), (e2, small, 0.5), (e1, big, 2.0)]):
            say(f"case densest_noalloc call {step}: n={mg.n} L={mg.L} beta={beta}")
            (ra, rd, rs, rv), (ga, gd, gs, gv) = both(eng, beta)
            assert ga == ra
            assert np.array_equal(gd, rd) and np.array_equal(gs, rs)
            assert np.array_equal(gv, rv)
            ow = O.wds(mg.n, mg.layers(), beta)
            assert (ga[0], ga[1], ga[2], ga[3]) == (ow["max_density"], ow["res_k"], ow["res_lmd"], ow["res_len"])


def test_noalloc_refuses_a_pass_above_the_array():
    """A 600-clique in 64 identical layers: kmax = 599, so the edge
    histograms alone take 8 * 64 * 600 bytes > 256 KiB."""
    n, L = 600, 64
    es = G.clique(range(n))
    mg = G.graph(n, [es] * L)
    assert 8 * L * n > SCRATCH_BYTES
    say("case densest_noalloc refusal (600-clique, 64 layers)")
    with Engine.from_graph(mg) as eng:
        eng.decompose()
        w = Wds()
        assert lib().hg_densest_noalloc(eng._h, C.c_double(1.0), C.byref(w)) == -4
        assert b"HG_NOALLOC_SCRATCH_BYTES" in lib().hg_last_error()
        ref = eng.densest(1.0)  # the allocating entry point still answers
        assert ref.res_k == n - 1
