"""GPU parity of the partitioned path (-a d, detectSCC): local SCCs equal
the oracle's SCCs of the union of partition-induced subgraphs (and the
known local-SCC counts of the reference's partition files), and the merged
labels equal Tarjan's on the whole graph."""
import subprocess

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


@pytest.mark.parametrize("idx", range(6))
def test_partition_fixtures(ctx, known, fixture_path, idx):
    pa = known["partitions"][idx]
    g = scc.load_full_graph(fixture_path(known["graphs"][pa["graph"]]["file"]))
    part = scc.load_partition(fixture_path(pa["part"]), g.V)
    ctx.upload(g)
    ctx.upload_partition(part)
    loc = ctx.phase_local()
    want_loc, n_loc = orc.local_scc(g, part)
    assert n_loc == pa["local_sccs"]
    assert_same_labels(loc, want_loc, f"local {pa['part']}")
    want, n = orc.tarjan(g)
    # trim-2 tests the live bitmap after the single-class trim-1 (-p 1 -q 1)
    # and the class words without it (-p 0 -q 1)
    for t1, t2 in [(1, 1), (0, 0), (1, 0), (0, 1)]:
        # the reference's vertex-per-lane mapping and the automatic expansion
        for ex in (scc.EXPAND_VERTEX, scc.EXPAND_AUTO):
            st = ctx.run(algo="d", trim1=t1, trim2=t2, expand=ex)
            assert_same_labels(ctx.labels(), want, f"detect {pa['part']} -p {t1} -q {t2} expand {ex}")
            assert st.num_sccs == n == pa["global_sccs"]


@pytest.mark.parametrize("seed", range(8))
def test_random_partitions(ctx, seed):
    rng = np.random.default_rng(300 + seed)
    V = int(rng.integers(2, 5000))
    E = int(rng.integers(0, 5 * V))
    src = rng.integers(1, V + 1, E)
    dst = rng.integers(1, V + 1, E)
    g = scc.from_arcs(V, src, dst)
    K = int(rng.integers(1, V + 1))
    part = np.zeros(V + 2, np.uint32)
    part[1:V + 1] = rng.integers(0, K, V) * 7 + 3          # sparse, non-dense ids
    ctx.upload(g)
    ctx.upload_partition(part)
    assert_same_labels(ctx.phase_local(), orc.local_scc(g, part)[0], f"local seed {seed}")
    want, n = orc.tarjan(g)
    for ex in (scc.EXPAND_VERTEX, scc.EXPAND_AUTO):
        st = ctx.run(algo="d", expand=ex)
        assert_same_labels(ctx.labels(), want, f"detect seed {seed} K={K} expand {ex}")
        assert st.num_sccs == n


def test_merge_quotient_shapes(ctx):
    """The merge's quotient graph at both extremes, in one context so its
    grow-only key buffer is reused, grown and then left larger than needed:
    far more quotient vertices than arcs (every vertex its own partition,
    trim-1 off, a few thousand arcs over 100,000 vertices), then far more
    arcs than quotient vertices (two partitions joined by many parallel
    arcs), then the first graph again."""
    rng = np.random.default_rng(77)
    V = 100_000
    src = rng.integers(1, V + 1, 3000)
    dst = rng.integers(1, V + 1, 3000)
    sparse = scc.from_arcs(V, np.r_[src, np.arange(1, 200)], np.r_[dst, np.arange(2, 201)])
    part_sparse = np.zeros(V + 2, np.uint32)
    part_sparse[1:V + 1] = np.arange(V, dtype=np.uint32)
    n2 = 4000
    a = rng.integers(1, n2 + 1, 60_000)
    b = rng.integers(1, n2 + 1, 60_000)
    dense = scc.from_arcs(n2, a, b)
    part_dense = np.zeros(n2 + 2, np.uint32)
    part_dense[1:n2 + 1] = (np.arange(1, n2 + 1) % 2).astype(np.uint32)
    for g, part, t1 in [(sparse, part_sparse, 0), (dense, part_dense, 1), (sparse, part_sparse, 0)]:
        ctx.upload(g)
        ctx.upload_partition(part)
        want, n = orc.tarjan(g)
        st = ctx.run(algo="d", trim1=t1, trim2=0)
        assert_same_labels(ctx.labels(), want, f"merge V={g.V} E={g.E} trim1={t1}")
        assert st.num_sccs == n


def test_planted_communities(ctx):
    g = scc.gen_communities(40000, 20, deg=8.0, inter=0.02, seed=55)
    ctx.upload(g)                       # uploads g.part too
    want, n = orc.tarjan(g)
    st = ctx.run(algo="d")
    assert_same_labels(ctx.labels(), want, "communities")
    assert st.num_sccs == n
    # the same graph with a seeded random partition stresses the merge
    rng = np.random.default_rng(5)
    part = np.zeros(g.V + 2, np.uint32)
    part[1:g.V + 1] = rng.integers(0, 200, g.V)
    ctx.upload_partition(part)
    ctx.run(algo="d")
    assert_same_labels(ctx.labels(), want, "communities, random partition")


def test_cli_detect(known, fixture_path):
    r = subprocess.run([scc.CLI_PATH, "-a", "d", "-p", "1", "-q", "1", "-w", "1", "-f",
                        fixture_path("sample5.gr"), "-P", fixture_path("MeTiSInput5.graph.part.5")],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "\nNumber Of Sccs : 3\nTime : " in r.stdout


def test_cli_labels_output(tmp_path, fixture_path, known):
    out = tmp_path / "labels.txt"
    r = subprocess.run([scc.CLI_PATH, "-a", "g", "-p", "1", "-q", "1", "-w", "1", "-f",
                        fixture_path("sample128.gr"), "-o", str(out)],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("numberOf1Sccs: ") and "Number Of Sccs : 13" in r.stdout
    lab = np.array([0] + [int(x) for x in out.read_text().split()], np.uint32)
    g = scc.load_full_graph(fixture_path("sample128.gr"))
    assert_same_labels(lab, orc.tarjan(g)[0], "cli -o")


def test_cli_reads_csr_cache(tmp_path, fixture_path):
    """`scc -f` takes the binary CSR cache as well as GTgraph text, with the
    same output lines."""
    g = scc.load_full_graph(fixture_path("sample128.gr"))
    p = tmp_path / "sample128.csr"
    scc.write_csr(str(p), g)
    outs = []
    for f in (fixture_path("sample128.gr"), str(p)):
        r = subprocess.run([scc.CLI_PATH, "-a", "g", "-p", "1", "-q", "1", "-w", "1", "-f", f],
                           stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs.append(r.stdout.split("Time")[0])
    assert outs[0] == outs[1] and "Number Of Sccs : 13" in outs[1]
