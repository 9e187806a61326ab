"""The allocation-free size histogram (scc_histogram_noalloc) through the C
ABI: on two calls in a row, the second on a smaller graph, it returns
exactly what scc_histogram returns, which the test also checks against the
histogram of Tarjan's labels."""
import ctypes

import numpy as np
import pytest

import detectingscc_amd as scc
from oracle import pyoracle as orc
from helpers import random_graph

pytestmark = pytest.mark.gpu


def _noalloc(ctx, cap):
    pairs = np.zeros(2 * max(cap, 1), np.uint64)
    n = ctypes.c_uint32(0)
    st = scc.lib().scc_histogram_noalloc(ctx._ctx, pairs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                         ctypes.c_uint32(cap), ctypes.byref(n))
    assert st == 0, scc.lib().scc_last_error().decode()
    k = min(n.value, cap)
    return [(int(pairs[2 * i]), int(pairs[2 * i + 1])) for i in range(k)], n.value


def _tarjan_pairs(g):
    lab, _ = orc.tarjan(g)
    sizes = np.bincount(lab[1:].astype(np.int64))
    hist = np.bincount(sizes[sizes > 0])
    return [(k, int(c)) for k, c in enumerate(hist) if c]


def test_noalloc_histogram_matches_allocating_one():
    rng = np.random.default_rng(4242)
    graphs = [scc.gen_rmat(18, seed=5),                   # many sizes, a giant SCC
              random_graph(3000, 4000, rng)]              # smaller: the second call
    with scc.Context(0) as ctx:
        for g in graphs:
            ctx.upload(g)
            ctx.run(algo="g", expand=scc.EXPAND_DO, sizes=False)
            want, _ = ctx.histogram()
            got, n = _noalloc(ctx, 4096)
            assert got == want and n == len(want)
            assert got == _tarjan_pairs(g)
            # a short output buffer gets the smallest sizes and the full count
            short, n2 = _noalloc(ctx, 2)
            assert short == want[:2] and n2 == len(want)
