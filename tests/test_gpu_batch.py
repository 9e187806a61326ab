"""GPU parity of the batched persistent form and the continuous-batching
server: every graph's labels equal Tarjan's on that graph alone, for each
configuration word (algorithm, trims, lanes, workgroup size), for the
one-graph-per-launch baseline, after autotuning, and for requests served in
a mix of concurrent arrivals (the last in its own file, collected late)."""
import numpy as np
import pytest

import detectingscc_amd as scc
from oracle import pyoracle as orc
from helpers import assert_same_labels, random_graph

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = scc.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def mix():
    b = scc.gen_batch_mix(96, 6, 11, 1.5, 14.0, seed=20261)
    want = [orc.tarjan(b.graph(i)) for i in range(b.num_graphs)]
    return b, want


def _edge_graphs():
    rng = np.random.default_rng(77)
    gs = [
        scc.from_arcs(1, [], []),                                   # one vertex, no arcs
        scc.from_arcs(1, [1], [1]),                                 # self-loop only
        scc.from_arcs(2, [1, 2], [2, 1]),                           # a 2-cycle
        scc.from_arcs(100, [], []),                                 # arcless
        scc.from_arcs(3000, np.arange(1, 3001), np.r_[np.arange(2, 3001), 1]),   # one long cycle
        scc.from_arcs(2000, np.arange(1, 2000), np.arange(2, 2001)),             # a path: all trim-1
        scc.from_arcs(5, [1, 1, 1, 2, 2, 3, 4, 5], [2, 2, 2, 1, 3, 4, 5, 3]),      # duplicate arcs
    ]
    k = 40                                                          # complete digraph
    s, d = np.meshgrid(np.arange(1, k + 1), np.arange(1, k + 1))
    gs.append(scc.from_arcs(k, s.ravel(), d.ravel()))
    gs.append(random_graph(70000, 140000, rng))                     # beyond 2^16 vertices
    gs.append(random_graph(4096, 4096 * 16, rng))                   # dense
    return gs


def _check_batch(ctx, b, want, what):
    labs, nscc = ctx.batch_labels()
    for i in range(b.num_graphs):
        assert_same_labels(labs[i], want[i][0], f"{what} graph {i} (V={b.V[i]}, E={b.E[i]})")
        assert nscc[i] == want[i][1], f"{what} graph {i}: {nscc[i]} SCCs, want {want[i][1]}"


CONFIGS = [
    dict(algo="g", trim1=True, trim2=True, lanes=1, block=256),
    dict(algo="g", trim1=True, trim2=True, lanes=4, block=1024),
    dict(algo="g", trim1=False, trim2=False, lanes=16, block=512),
    dict(algo="g", trim1=True, trim2=False, lanes=64, block=256),
    dict(algo="h", trim1=True, trim2=False, lanes=4, block=256),
    dict(algo="h", trim1=False, trim2=False, lanes=1, block=1024),
]


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: "{algo}-t{trim1:d}{trim2:d}-l{lanes}-b{block}".format(**c))
def test_batch_mix(ctx, mix, cfg):
    b, want = mix
    ctx.batch_upload(b)
    st = ctx.batch_run(**cfg)
    assert st.graphs == b.num_graphs and st.launches == 1
    _check_batch(ctx, b, want, str(cfg))


def test_batch_edge_graphs(ctx):
    gs = _edge_graphs()
    b = scc.batch_from_graphs(gs)
    want = [orc.tarjan(g) for g in gs]
    ctx.batch_upload(b)
    for cfg in (CONFIGS[0], CONFIGS[2], CONFIGS[4]):
        ctx.batch_run(**cfg)
        _check_batch(ctx, b, want, f"edge {cfg}")


TIER_SIZES = (8190, 8193, 16383, 16386, 32767, 33000, 40000, 65000, 65536)


def _tier_graphs(dags):
    """Graphs at and around the LDS tier boundaries of every workgroup size
    (class words in LDS up to 2^13, 2^14 or 2^15 vertices, visited bitmaps
    up to 2^15 or 2^16): ER with planted cycles (non-trivial SCCs), or
    layered DAGs (all singletons; 8 layers, arcs only to later layers, so
    trim-1 peels one in at most 8 passes -- a deep DAG makes the colouring
    of vSlota quadratic, as in the reference)."""
    rng = np.random.default_rng(2027)
    gs = []
    for V in TIER_SIZES:
        if not dags:
            gs.append(scc.gen_er(V, 3 * V, cycles=8, cmin=2, cmax=512, seed=V))
            continue
        src = rng.integers(1, V + 1, 4 * V)
        dst = rng.integers(1, V + 1, 4 * V)
        layer = lambda x: (x - 1) * 8 // V
        keep = layer(src) < layer(dst)
        gs.append(scc.from_arcs(V, src[keep], dst[keep]))
    return gs


@pytest.mark.parametrize("block", [256, 512, 1024])
def test_batch_lds_tiers(ctx, block):
    for dags, algos in ((False, ("g", "h")), (True, ("g",))):
        gs = _tier_graphs(dags)
        b = scc.batch_from_graphs(gs)
        want = [orc.tarjan(g) for g in gs]
        ctx.batch_upload(b)
        for algo in algos:
            ctx.batch_run(algo=algo, trim1=True, trim2=algo == "g", lanes=16, block=block)
            _check_batch(ctx, b, want, f"tiers block {block} -a {algo} dags={dags}")
    gs = _tier_graphs(False)
    want = [orc.tarjan(g) for g in gs]
    # the single-graph engine agrees graph by graph
    with scc.Context(0) as one:
        for g, (lab, n) in zip(gs[:4], want[:4]):
            one.upload(g)
            st = one.run(algo="g")
            assert st.num_sccs == n and np.array_equal(one.labels(), lab)


def test_batch_serial_baseline(ctx, mix):
    b, want = mix
    ctx.batch_upload(b)
    st = ctx.batch_run_serial()
    assert st.graphs == b.num_graphs
    _check_batch(ctx, b, want, "one graph per launch")


def test_batch_autotune(ctx, mix):
    b, want = mix
    ctx.batch_upload(b)
    report = ctx.batch_autotune(sample=3)
    assert "bucket" in report
    st = ctx.batch_run(autotuned=True)
    assert 1 <= st.launches <= 3
    _check_batch(ctx, b, want, "autotuned")


def test_batch_reupload_grow_and_shrink(ctx, mix):
    b, want = mix
    small = scc.batch_from_graphs([b.graph(i) for i in range(5)])
    big = scc.gen_batch_mix(24, 10, 13, 2.0, 10.0, seed=5)
    big_want = [orc.tarjan(big.graph(i)) for i in range(big.num_graphs)]
    for bb, ww in ((small, want[:5]), (big, big_want), (small, want[:5])):
        ctx.batch_upload(bb)
        ctx.batch_run(**CONFIGS[1])
        _check_batch(ctx, bb, ww, f"reupload {bb.num_graphs}")


def test_batch_repeatable(ctx, mix):
    b, _ = mix
    ctx.batch_upload(b)
    ctx.batch_run(**CONFIGS[3])
    first = np.concatenate(ctx.batch_labels()[0])
    ctx.batch_run(**CONFIGS[3])
    assert np.array_equal(first, np.concatenate(ctx.batch_labels()[0]))


def test_server_budget_admission():
    """A tight budget splits the queue into many small batches; a request
    larger than the whole budget is rejected; results stay exact."""
    b = scc.gen_batch_mix(40, 8, 10, 2.0, 8.0, seed=41)
    graphs = [b.graph(i) for i in range(b.num_graphs)]
    def footprint(g):   # device bytes a request holds while admitted
        slots = ((g.V + 2 + 63) // 64) * 64
        return slots * 64 + slots // 4 + 8 * g.E + 48
    budget = 3 * max(footprint(g) for g in graphs)
    with scc.Context(0) as c:
        c.server_start(budget_bytes=budget, max_batch=64, slots=2)
        ids = [c.server_submit(g) for g in graphs]
        big = random_graph(200000, 400000, np.random.default_rng(1))
        with pytest.raises(scc.SCCError) as e:
            c.server_submit(big)
        assert e.value.status == 6
        c.server_wait_all()
        for g, rid in zip(graphs, ids):
            lab, n, _ = c.server_poll(rid)
            assert_same_labels(lab, orc.tarjan(g)[0], f"budgeted request {rid}")
        s = c.server_stats()
        assert s.rejected == 1 and s.completed == len(graphs)
        assert s.batches >= -(-sum(footprint(g) for g in graphs) // budget)
        assert s.peak_bytes <= budget
        c.server_stop()
