"""GPU parity above 1M vertices, where the sweeps run 2,048 workgroups (every
wave slot of the chip): spread counters and slotted trim-1 lists and
bottom-up queues under that many workgroups, the first-neighbour probes,
and both label paths of Hong's first round (from the search's bitmaps
without sizes, eager with sizes), against Tarjan."""
import numpy as np
import pytest

import detectingscc_amd as scc
from oracle import pyoracle as orc
from helpers import assert_same_labels

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rmat21():
    g = scc.gen_rmat(21, seed=21)   # 2,097,152 V, 33,554,432 arcs
    want, n = orc.tarjan(g)
    return g, want, n


def test_rmat21_hong_both_label_paths(rmat21):
    g, want, n = rmat21
    with scc.Context(0) as ctx:
        ctx.upload(g)
        for sizes in (False, True, False):
            st = ctx.run(algo="g", expand=scc.EXPAND_DO, sizes=sizes)
            assert_same_labels(ctx.labels(), want, f"rmat21 -a g sizes={sizes}")
            assert st.num_sccs == n
            if sizes:
                assert st.largest == int(np.bincount(want[1:].astype(np.int64)).max())
        st = ctx.run(algo="g", expand=scc.EXPAND_DO, sizes=False, phase_times=True)
        assert_same_labels(ctx.labels(), want, "rmat21 -a g phase times")
        assert st.ms_phase["reach"] > 0 and st.ms_phase["trim1"] > 0


def test_rmat21_slota_and_vertex_mapping(rmat21):
    g, want, n = rmat21
    with scc.Context(0) as ctx:
        ctx.upload(g)
        for algo, expand in (("h", scc.EXPAND_DO), ("g", scc.EXPAND_VERTEX)):
            st = ctx.run(algo=algo, expand=expand, sizes=False)
            assert_same_labels(ctx.labels(), want, f"rmat21 -a {algo} expand {expand}")
            assert st.num_sccs == n


def test_one_context_many_sizes():
    """Graphs of decreasing and increasing size in one context: the queues'
    slot regions follow each graph's grid sizes and must fit the buffers
    the context keeps (sizes around the 1,024-workgroup boundary)."""
    graphs = [scc.gen_er(300000, 600000, cycles=10, cmin=2, cmax=500, seed=1),
              scc.gen_er(262143, 520000, cycles=10, cmin=2, cmax=500, seed=2),
              scc.gen_rmat(14, seed=5),
              scc.gen_er(65537, 200000, cycles=10, cmin=2, cmax=500, seed=3),
              scc.gen_er(262145, 520000, cycles=10, cmin=2, cmax=500, seed=4)]
    with scc.Context(0) as ctx:
        for i, g in enumerate(graphs):
            want, n = orc.tarjan(g)
            ctx.upload(g)
            for sizes in (False, True):
                st = ctx.run(algo="g", expand=scc.EXPAND_DO, sizes=sizes)
                assert_same_labels(ctx.labels(), want, f"graph {i} V={g.V} sizes={sizes}")
                assert st.num_sccs == n
