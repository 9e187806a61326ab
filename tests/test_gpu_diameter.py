"""GPU parity on high-diameter and mixed graphs in direction-optimising
mode, where several top-down levels are queued per host check: Tarjan's
labels for vHong and vSlota, and the BFS reach sets in both directions with
and without a dead set."""
import numpy as np
import pytest

import detectingscc_amd as scc
from oracle import pyoracle as orc
from helpers import assert_same_labels

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = scc.Context(0)
    yield c
    c.close()


def _cycle(n):
    ids = np.arange(1, n + 1)
    return scc.from_arcs(n, ids, np.r_[ids[1:], 1])


def _ladder(n):
    """Two opposite paths joined every 50 rungs: one SCC, diameter ~n."""
    a = np.arange(1, n)
    src = np.r_[a, a + n + 1, np.arange(1, n + 1, 50)]
    dst = np.r_[a + 1, a + n, np.arange(1, n + 1, 50) + n]
    src = np.r_[src, np.arange(n + 1, 2 * n + 1, 50)]
    dst = np.r_[dst, np.arange(1, n + 1, 50)]
    return scc.from_arcs(2 * n, src, dst)


def _barbell(na, nb, pairs, seed):
    """Dense core A -> a chain of `pairs` 2-cycles -> denser core B.  The
    backward search from B's hub is wide, then one pair wide for `pairs`
    levels, then wide again in A, so the search switches direction on both
    sides of a long run of narrow levels."""
    rng = np.random.default_rng(seed)

    def core(base, n, deg):
        s = rng.integers(0, n, size=n * deg)
        t = rng.integers(0, n, size=n * deg)
        ring = np.arange(n)
        return np.r_[s, ring] + base, np.r_[t, (ring + 1) % n] + base

    sa, ta = core(1, na, 8)
    sb, tb = core(1 + na + 2 * pairs, nb, 16)
    a = 1 + na + 2 * np.arange(pairs)       # first vertex of each pair
    src = np.r_[sa, sb, a, a + 1, a[:-1], [na], [a[-1]]]
    dst = np.r_[ta, tb, a + 1, a, a[1:], [a[0]], [1 + na + 2 * pairs]]
    return scc.from_arcs(na + 2 * pairs + nb, src, dst)


GRAPHS = {
    "grid256": lambda: scc.gen_grid(256, 0.7, seed=4096),
    "barbell": lambda: _barbell(20000, 30000, 3000, seed=11),
    "grid512": lambda: scc.gen_grid(512, 0.75, seed=7),
    "cycle20000": lambda: _cycle(20000),
    "ladder": lambda: _ladder(6000),
    "rmat14": lambda: scc.gen_rmat(14, seed=3),
    "er_planted": lambda: scc.gen_er(30000, 60000, cycles=40, cmin=2, cmax=3000, seed=9),
}


@pytest.mark.parametrize("name", list(GRAPHS))
def test_high_diameter_matches_tarjan(ctx, name):
    g = GRAPHS[name]()
    want, n = orc.tarjan(g)
    ctx.upload(g)
    for algo in ("g", "h"):
        st = ctx.run(algo=algo, expand=scc.EXPAND_DO, sizes=False)
        assert_same_labels(ctx.labels(), want, f"{name} -a {algo}")
        assert st.num_sccs == n, (name, algo)


@pytest.mark.parametrize("name", ["grid512", "ladder", "rmat14", "barbell"])
def test_high_diameter_reach_sets(ctx, name):
    g = GRAPHS[name]()
    ctx.upload(g)
    rng = np.random.default_rng(5)
    srcs = np.r_[rng.integers(1, g.V + 1, size=3), g.V]
    dead = (np.arange(g.V + 1) % 13 == 5).astype(np.uint8)
    for src in srcs:
        dead[src] = 0
        for d in (0, 1):
            for dd in (None, dead):
                want = orc.reach(g, int(src), d, dead=dd)
                got, levels, _ = ctx.phase_reach(int(src), d, scc.EXPAND_DO, 1, dead=dd)
                assert np.array_equal(got, want), (name, int(src), d, dd is None)
