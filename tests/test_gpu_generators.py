"""Generators on the device: the graph a context generates in place is
bit-identical (CSR and reverse CSR, and the partition of the community
graph) to the host generator's with the same arguments, at the benchmark
configurations' sizes."""
import numpy as np
import pytest

import detectingscc_amd as scc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = scc.Context(0)
    yield c
    c.close()


def _same(dev, host, what):
    assert (dev.V, dev.E) == (host.V, host.E), what
    for a in ("Fr", "Fc", "Br", "Bc"):
        assert np.array_equal(getattr(dev, a), getattr(host, a)), f"{what}: {a} differs"


def test_er_planted_cycles(ctx):
    ctx.gen_er_dev(16384, 131072, cycles=64, cmin=2, cmax=256, seed=1)
    _same(ctx.download(), scc.gen_er(16384, 131072, cycles=64, cmin=2, cmax=256, seed=1), "ER 16,384 + 64 cycles")


@pytest.mark.parametrize("V,draws,cycles,cmax", [(1, 10, 3, 256), (5, 0, 4, 3), (3000, 9000, 200, 512)])
def test_er_edge_cases(ctx, V, draws, cycles, cmax):
    ctx.gen_er_dev(V, draws, cycles=cycles, cmin=2, cmax=cmax, seed=9)
    _same(ctx.download(), scc.gen_er(V, draws, cycles=cycles, cmin=2, cmax=cmax, seed=9), f"ER {V}")


@pytest.mark.parametrize("scale,seed", [(10, 1), (22, 22)])
def test_rmat(ctx, scale, seed):
    ctx.gen_rmat_dev(scale, seed=seed)
    _same(ctx.download(), scc.gen_rmat(scale, seed=seed), f"R-MAT {scale}")


def test_rmat24_rows(ctx):
    """R-MAT scale 24 (268M arcs): the full arrays."""
    ctx.gen_rmat_dev(24, seed=24)
    _same(ctx.download(), scc.gen_rmat(24, seed=24), "R-MAT 24")


@pytest.mark.parametrize("side", [1, 7, 4096])
def test_grid(ctx, side):
    ctx.gen_grid_dev(side, 0.7, seed=4096)
    _same(ctx.download(), scc.gen_grid(side, 0.7, seed=4096), f"grid {side}")


def test_communities(ctx):
    ctx.gen_communities_dev(1 << 23, 200, 12.0, 0.02, seed=55)
    host = scc.gen_communities(1 << 23, 200, 12.0, 0.02, seed=55)
    _same(ctx.download(), host, "communities")
    assert np.array_equal(ctx.graph.part, host.part)


def test_device_graph_decomposes(ctx):
    """A device-generated graph goes straight into a decomposition."""
    from oracle import pyoracle as orc
    ctx.gen_rmat_dev(16, seed=16)
    st = ctx.run(algo="g", expand=scc.EXPAND_DO)
    want, n = orc.tarjan(scc.gen_rmat(16, seed=16))
    assert st.num_sccs == n and np.array_equal(ctx.labels(), want)


@pytest.mark.parametrize("graphs,lgmin,lgmax,seed", [(96, 6, 11, 20261), (4096, 8, 16, 4096)])
def test_batch_mix(ctx, graphs, lgmin, lgmax, seed):
    """The batched mix (40% ER, 40% R-MAT, 20% planted cycles + ER fill)
    generated on the device equals the host generator's, array by array,
    and the device-built batch decomposes like the host-built one."""
    dev = ctx.gen_batch_mix_dev(graphs, lgmin, lgmax, 2.0 if graphs == 4096 else 1.5,
                                16.0 if graphs == 4096 else 14.0, seed=seed, copy=True)
    host = scc.gen_batch_mix(graphs, lgmin, lgmax, 2.0 if graphs == 4096 else 1.5,
                             16.0 if graphs == 4096 else 14.0, seed=seed)
    for a in ("V", "E", "row_off", "col_off", "Fr", "Fc", "Br", "Bc"):
        assert np.array_equal(getattr(dev, a), getattr(host, a)), f"{graphs}-graph mix: {a} differs"
    if graphs == 96:
        from oracle import pyoracle as orc
        ctx.batch_run(algo="g", lanes=16, block=1024)
        labs, nscc = ctx.batch_labels()
        for i in range(graphs):
            want, n = orc.tarjan(host.graph(i))
            assert nscc[i] == n and np.array_equal(labs[i], want), f"graph {i}"
