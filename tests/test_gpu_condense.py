"""GPU parity of the condensation DAG: nodes are the SCCs in order of their
canonical (minimum-id) labels, arcs the distinct pairs of SCCs joined by an
arc.  Checked against the DAG built from Tarjan's labels on the CPU, and by
the defining property that every SCC of the DAG is a single node."""
import subprocess

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


def _dag_from_labels(g, lab):
    roots = np.unique(lab[1:])
    node = np.zeros(g.V + 1, np.uint32)
    node[roots] = np.arange(1, roots.size + 1, dtype=np.uint32)
    src = np.repeat(np.arange(g.V + 1, dtype=np.int64), np.diff(g.Fr.astype(np.int64)))
    a = node[lab[src]]
    b = node[lab[g.Fc]]
    keep = a != b
    pairs = np.unique(np.stack([a[keep], b[keep]], axis=1), axis=0) if keep.any() else np.zeros((0, 2), np.uint32)
    dag = scc.from_arcs(int(roots.size), pairs[:, 0].astype(np.uint32), pairs[:, 1].astype(np.uint32))
    return dag, node[lab]


def _check(ctx, g, what):
    want_lab, _ = orc.tarjan(g)
    dag, node_of = ctx.condensation()
    want_dag, want_node = _dag_from_labels(g, want_lab)
    assert (dag.V, dag.E) == (want_dag.V, want_dag.E), what
    for a in ("Fr", "Fc", "Br", "Bc"):
        assert np.array_equal(getattr(dag, a), getattr(want_dag, a)), f"{what}: {a}"
    assert np.array_equal(node_of[1:], want_node[1:]), what
    _, n = orc.tarjan(dag)
    assert n == dag.V, f"{what}: the condensation has a cycle"


GRAPHS = {
    "sample128": lambda fp: scc.load_full_graph(fp("sample128.gr")),
    "sample5": lambda fp: scc.load_full_graph(fp("sample5.gr")),
    "er": lambda fp: random_graph(3000, 4500, np.random.default_rng(2)),
    "rmat12": lambda fp: scc.gen_rmat(12, seed=8),
    "planted": lambda fp: scc.gen_er(5000, 6000, cycles=30, cmin=2, cmax=100, seed=3),
    "one_scc": lambda fp: scc.from_arcs(4, [1, 2, 3, 4], [2, 3, 4, 1]),
    "arcless": lambda fp: scc.from_arcs(7, [], []),
}


@pytest.mark.parametrize("name", list(GRAPHS))
@pytest.mark.parametrize("algo", ["g", "h"])
def test_condensation(ctx, fixture_path, name, algo):
    g = GRAPHS[name](fixture_path)
    ctx.upload(g)
    ctx.run(algo=algo, expand=scc.EXPAND_DO if algo == "g" else scc.EXPAND_AUTO, sizes=False)
    _check(ctx, g, f"{name} -a {algo}")


def test_condensation_after_detect(ctx, known, fixture_path):
    pa = known["partitions"][3]
    g = scc.load_full_graph(fixture_path(known["graphs"][pa["graph"]]["file"]))
    ctx.upload(g)
    ctx.upload_partition(scc.load_partition(fixture_path(pa["part"]), g.V))
    ctx.run(algo="d")
    _check(ctx, g, "detectSCC")


def test_condensation_needs_a_run():
    with scc.Context(0) as c:
        c.upload(scc.from_arcs(3, [1], [2]))
        with pytest.raises(scc.SCCError) as e:
            c.condensation()
        assert e.value.status == 8


def test_cli_writes_condensation(tmp_path, fixture_path):
    out = tmp_path / "dag.gr"
    r = subprocess.run([scc.CLI_PATH, "-a", "g", "-p", "1", "-q", "1", "-w", "1", "-f",
                        fixture_path("sample128.gr"), "-c", str(out)],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Number Of Sccs : 13" in r.stdout
    dag = scc.load_full_graph(str(out))
    g = scc.load_full_graph(fixture_path("sample128.gr"))
    want, _ = _dag_from_labels(g, orc.tarjan(g)[0])
    assert dag.V == 13 and np.array_equal(dag.Fc, want.Fc) and np.array_equal(dag.Fr, want.Fr)
