"""GPU parity at the five benchmark configurations' own sizes: every run's
canonical labels, SCC count, largest SCC and size histogram equal host
Tarjan's on the same seeded graph.

  ER 16,384 V / 131,072 draws + 64 planted cycles (2..256), seed 1 (the
      small vHong case; also in test_gpu_parity.py)
  R-MAT scale 22 (4,194,304 V, 67,108,864 arcs), seed 22: all 64 runs of
      {g, h} x (p, q) in {0,1}^2 and {x, y} x W in {1..64} x (p, q) in {0,1}^2
      (src/scc.cu:78-102 dispatches the four drivers)
  R-MAT scale 24 (16,777,216 V, 268,435,456 arcs), seed 24: -a g and -a h
  grid 4096 x 4096, arc probability 0.7, seed 4096: -a g, -a x, -a h
  the 4,096-graph batch mix (V log-uniform in [2^8, 2^16], mean degree
      uniform in [2, 16], 40% ER / 40% R-MAT / 20% planted cycles), seed
      4096: every graph, batched (autotuned and a fixed configuration) and
      one graph per launch
  serving: 400 lognormal requests (median 2^14 V, sigma 1.5, capped at
      2^22, degree 8), seed 8, with requests of V >= 2^16 on the device-wide
      engine; and the 8,388,608-vertex graph with 200 planted communities
      (mean degree 12, 2% inter-community arcs), seed 55, through -a d with
      its 200 community partitions and with 200 random partitions
      (src/scc_WCC.cu:1000-1617)
"""
import numpy as np
import pytest

import detectingscc_amd as scc
from oracle import pyoracle as orc
from helpers import assert_same_labels

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

WARPS = [1, 2, 4, 8, 16, 32, 64]
TRIMS = [(1, 1), (1, 0), (0, 1), (0, 0)]


def _hist(lab):
    sizes = np.bincount(lab[1:].astype(np.int64))
    sizes = sizes[sizes > 0]
    return np.bincount(sizes)


class Truth:
    """Tarjan's labels of a graph, with the count, largest SCC and the size
    histogram (number of SCCs of each size)."""

    def __init__(self, g):
        self.lab, self.n = orc.tarjan(g)
        self.hist = _hist(self.lab)
        self.largest = len(self.hist) - 1

    def pairs(self):
        return [(k, int(c)) for k, c in enumerate(self.hist) if c]

    def check(self, lab, st, what, sized=True, ctx=None):
        assert_same_labels(lab, self.lab, what)
        assert st.num_sccs == self.n, f"{what}: {st.num_sccs} SCCs, want {self.n}"
        h = _hist(lab)
        assert np.array_equal(h, self.hist), f"{what}: size histogram differs"
        if sized:
            assert st.largest == self.largest, f"{what}: largest {st.largest}, want {self.largest}"
        if ctx is not None:   # the histogram computed on the device
            dh, _ = ctx.histogram()
            assert dh == self.pairs(), f"{what}: device size histogram differs"


@pytest.fixture(scope="module")
def ctx():
    c = scc.Context(0)
    yield c
    c.close()


# ---------------------------------------------------------------- C1 -------

def test_er16k_planted_cycles(ctx):
    g = scc.gen_er(16384, 131072, cycles=64, cmin=2, cmax=256, seed=1)
    t = Truth(g)
    ctx.upload(g)
    for algo, w in [("g", 1), ("h", 1), ("x", 32), ("y", 64)]:
        for p, q in TRIMS:
            st = ctx.run(algo=algo, warp=w, trim1=p, trim2=q)
            t.check(ctx.labels(), st, f"ER16k -a {algo} -w {w} -p {p} -q {q}", ctx=ctx)


# ---------------------------------------------------------------- C2 -------

@pytest.fixture(scope="module")
def rmat22(ctx):
    g = scc.gen_rmat(22, seed=22)
    t = Truth(g)
    ctx.upload(g)
    return g, t


C2_RUNS = ([(a, 1, p, q) for a in "gh" for p, q in TRIMS] +
           [(a, w, p, q) for a in "xy" for w in WARPS for p, q in TRIMS])
assert len(C2_RUNS) == 64


def ref_mapping(algo):
    """The reference's own frontier mapping for an algorithm: a W-lane
    virtual warp for -a x / -a y, a vertex per lane otherwise."""
    return scc.EXPAND_VWARP if algo in "xy" else scc.EXPAND_VERTEX


@pytest.mark.parametrize("algo,warp,p,q", C2_RUNS)
def test_rmat22_matrix(ctx, rmat22, algo, warp, p, q):
    g, t = rmat22
    if ctx.graph is not g:
        ctx.upload(g)
    st = ctx.run(algo=algo, warp=warp, trim1=p, trim2=q, expand=ref_mapping(algo))
    t.check(ctx.labels(), st, f"R-MAT 22 -a {algo} -w {warp} -p {p} -q {q}", ctx=ctx)


@pytest.mark.parametrize("algo", "ghxy")
@pytest.mark.parametrize("p,q", TRIMS)
def test_rmat22_default_expansion(ctx, rmat22, algo, p, q):
    """What a drop-in caller gets: the automatic expansion (direction-optimised)."""
    g, t = rmat22
    if ctx.graph is not g:
        ctx.upload(g)
    st = ctx.run(algo=algo, warp=8, trim1=p, trim2=q, expand=scc.EXPAND_AUTO)
    t.check(ctx.labels(), st, f"R-MAT 22 -a {algo} -p {p} -q {q} default expansion", ctx=ctx)


def test_rmat22_direction_optimised_unsized(ctx, rmat22):
    """The benchmark's timed form: direction-optimised expansion without the
    sizes pass (labels from the first search's bitmaps)."""
    g, t = rmat22
    if ctx.graph is not g:
        ctx.upload(g)
    for _ in range(2):
        st = ctx.run(algo="g", expand=scc.EXPAND_DO, sizes=False)
        t.check(ctx.labels(), st, "R-MAT 22 -a g DO unsized", sized=False, ctx=ctx)


# ---------------------------------------------------------------- C3 -------

def test_rmat24(ctx):
    g = scc.gen_rmat(24, seed=24)
    t = Truth(g)
    ctx.upload(g)
    for algo, w, expand, sizes in [("g", 1, scc.EXPAND_DO, False), ("g", 1, scc.EXPAND_VERTEX, True),
                                   ("h", 1, scc.EXPAND_VERTEX, True), ("x", 32, scc.EXPAND_VWARP, True),
                                   ("h", 1, scc.EXPAND_AUTO, True), ("y", 8, scc.EXPAND_AUTO, False)]:
        st = ctx.run(algo=algo, warp=w, expand=expand, sizes=sizes)
        t.check(ctx.labels(), st, f"R-MAT 24 -a {algo} -w {w} expand {expand}", sized=sizes, ctx=ctx)


def test_grid4096(ctx):
    g = scc.gen_grid(4096, 0.7, seed=4096)
    t = Truth(g)
    ctx.upload(g)
    for algo, w, expand in [("g", 1, scc.EXPAND_DO), ("g", 1, scc.EXPAND_VERTEX), ("x", 4, scc.EXPAND_VWARP),
                            ("h", 1, scc.EXPAND_VERTEX), ("h", 1, scc.EXPAND_AUTO)]:
        st = ctx.run(algo=algo, warp=w, expand=expand)
        t.check(ctx.labels(), st, f"grid 4096 -a {algo} -w {w} expand {expand}", ctx=ctx)


# ---------------------------------------------------------------- C4 -------

def test_batch_mix4096(ctx):
    b = scc.gen_batch_mix(4096, 8, 16, 2.0, 16.0, seed=4096)
    V = b.V.astype(np.int64)
    assert ((V > (1 << 15)) & (V <= (1 << 16))).sum() > 100    # the HBM-class-word tier is in the mix
    want = [orc.tarjan(b.graph(i)) for i in range(b.num_graphs)]
    ctx.batch_upload(b)

    def check(what):
        labs, nscc = ctx.batch_labels()
        for i in range(b.num_graphs):
            assert_same_labels(labs[i], want[i][0], f"{what} graph {i} (V={b.V[i]}, E={b.E[i]})")
            assert nscc[i] == want[i][1], f"{what} graph {i}: {nscc[i]} SCCs, want {want[i][1]}"

    ctx.batch_autotune(sample=256)
    ctx.batch_run(autotuned=True)
    check("batched, autotuned")
    ctx.batch_run(algo="g", trim1=True, trim2=True, lanes=16, block=1024)
    check("batched, Hong lanes 16 block 1024")
    ctx.batch_run_serial(expand=scc.EXPAND_DO)
    check("one graph per launch")


# ---------------------------------------------------------------- C5 -------

def test_serving_mix(ctx):
    b = scc.gen_serving_mix(400, 16384, 1.5, 1 << 22, 8.0, seed=8)
    graphs = [b.graph(i) for i in range(b.num_graphs)]
    assert max(g.V for g in graphs) >= 1 << 20 and sum(g.V >= 1 << 16 for g in graphs) >= 20
    ctx.server_start(budget_bytes=32 << 30, max_batch=256, slots=3, solo_vertices=1 << 16, lanes=16,
                     block=1024)
    try:
        ids = [ctx.server_submit(g) for g in graphs]
        ctx.server_wait_all()
        for i, (g, rid) in enumerate(zip(graphs, ids)):
            lab, n, _ = ctx.server_poll(rid)
            want, wn = orc.tarjan(g)
            assert_same_labels(lab, want, f"request {i} (V={g.V}, E={g.E})")
            assert n == wn, f"request {i}: {n} SCCs, want {wn}"
        st = ctx.server_stats()
        assert st.completed == len(graphs) and st.failed == 0 and st.solo >= 20
    finally:
        ctx.server_stop()


def test_communities_partitioned(ctx):
    g = scc.gen_communities(1 << 23, 200, 12.0, 0.02, seed=55)
    assert len(np.unique(g.part[1:g.V + 1])) == 200
    t = Truth(g)
    ctx.upload(g)                      # uploads the community partition too
    loc_want, n_loc = orc.local_scc(g, g.part)
    loc = ctx.phase_local()
    assert_same_labels(loc, loc_want, "local SCCs, 200 community partitions")
    for p, q in [(1, 1), (0, 0)]:
        st = ctx.run(algo="d", trim1=p, trim2=q)
        t.check(ctx.labels(), st, f"-a d, 200 community partitions, -p {p} -q {q}", ctx=ctx)
    rng = np.random.default_rng(200)
    part = np.zeros(g.V + 2, np.uint32)
    part[1:g.V + 1] = rng.integers(0, 200, g.V)
    ctx.upload_partition(part)
    assert_same_labels(ctx.phase_local(), orc.local_scc(g, part)[0], "local SCCs, 200 random partitions")
    st = ctx.run(algo="d")
    t.check(ctx.labels(), st, "-a d, 200 random partitions")
