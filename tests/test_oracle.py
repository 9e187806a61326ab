"""Pins the CPU oracle before anything is checked against it.

- Appendix-C known answers of the reference's sample inputs (counts, size
  classes, largest SCC, label digest) computed independently of this code.
- Three independent SCC algorithms agree: Tarjan, Kosaraju, brute-force
  transitive closure, exhaustively on every digraph with n <= 3 (self-loops
  included) and on a sample of n = 4, plus random graphs up to n = 63.
- Closed forms of the phase restatements (trim-1 on a DAG removes all,
  removed vertices are singletons, FW(p) & BW(p) = SCC(p), min/max
  reachable id by brute force, planted cycles are SCCs).
- The CPU ports of vHong / vSlota equal Tarjan.
"""
import itertools

import numpy as np
import pytest

from oracle import pyoracle as orc
from helpers import digest, sizes_of


def _load(known, fixture_path, name):
    return orc.load_gtgraph(fixture_path(known["graphs"][name]["file"]))


@pytest.mark.parametrize("name", ["sample0", "sample4", "sample5", "sample1", "sample128", "sample2"])
def test_oracle_known_answers(known, fixture_path, name):
    ka = known["graphs"][name]
    g = _load(known, fixture_path, name)
    assert (g.V, g.E) == (ka["V"], ka["E"])
    arcs = np.stack([np.repeat(np.arange(g.V + 1), np.diff(g.Fr.astype(np.int64))), g.Fc.astype(np.int64)])
    assert arcs.shape[1] - np.unique(arcs, axis=1).shape[1] == ka["dup_arcs"]
    lab, n = orc.tarjan(g)
    s = sizes_of(lab)
    assert n == ka["sccs"] and len(s) == ka["sccs"]
    assert (s == 1).sum() == ka["size1"] and (s == 2).sum() == ka["size2"]
    assert s.max() == ka["largest"]
    assert digest(lab) == ka["digest"]


def test_oracle_digest_rejects_planted_mistake(known, fixture_path):
    """A single wrong label changes the digest (the check is not vacuous)."""
    g = _load(known, fixture_path, "sample128")
    lab, _ = orc.tarjan(g)
    bad = lab.copy()
    v = int(np.flatnonzero(lab[1:] != np.arange(1, g.V + 1))[0]) + 1   # a non-root vertex
    bad[v] = v
    assert digest(bad) != known["graphs"]["sample128"]["digest"]


def _graph_from_bits(n, bits):
    src, dst = [], []
    for k, (u, v) in enumerate(itertools.product(range(1, n + 1), repeat=2)):
        if bits >> k & 1:
            src.append(u)
            dst.append(v)
    return orc.build_csr(n, src, dst)


@pytest.mark.parametrize("n", [1, 2, 3])
def test_three_algorithms_agree_exhaustive(n):
    for bits in range(1 << (n * n)):
        g = _graph_from_bits(n, bits)
        a, na = orc.tarjan(g)
        b, nb = orc.kosaraju(g)
        c, nc = orc.brute(g)
        assert na == nb == nc and (a == b).all() and (a == c).all(), (n, bits)


def test_three_algorithms_agree_n4_sample():
    rng = np.random.default_rng(4)
    for bits in rng.choice(1 << 16, size=3000, replace=False):
        g = _graph_from_bits(4, int(bits))
        a, na = orc.tarjan(g)
        c, nc = orc.brute(g)
        b, nb = orc.kosaraju(g)
        assert na == nb == nc and (a == b).all() and (a == c).all(), int(bits)


@pytest.mark.parametrize("seed", range(40))
def test_three_algorithms_agree_random(seed):
    rng = np.random.default_rng(seed)
    V = int(rng.integers(5, 64))
    E = int(rng.integers(0, 4 * V))
    g = orc.build_csr(V, rng.integers(1, V + 1, E), rng.integers(1, V + 1, E))
    a, na = orc.tarjan(g)
    b, nb = orc.kosaraju(g)
    c, nc = orc.brute(g)
    assert na == nb == nc and (a == b).all() and (a == c).all()


def _reach_matrix(g):
    V = g.V
    R = np.eye(V + 1, dtype=bool)
    src = np.repeat(np.arange(V + 1), np.diff(g.Fr.astype(np.int64)))
    R[src, g.Fc] = True
    for k in range(V + 1):
        R |= R[:, [k]] & R[[k], :]
    return R


@pytest.mark.parametrize("seed", range(20))
def test_phase_closed_forms_small(seed):
    rng = np.random.default_rng(100 + seed)
    V = int(rng.integers(2, 40))
    E = int(rng.integers(0, 3 * V))
    g = orc.build_csr(V, rng.integers(1, V + 1, E), rng.integers(1, V + 1, E))
    lab, _ = orc.tarjan(g)
    R = _reach_matrix(g)
    # min / max reachable id by brute force
    ids = np.arange(V + 1)
    want_min = np.array([0] + [ids[1:][R[v, 1:]].min() for v in range(1, V + 1)])
    want_max = np.array([0] + [ids[1:][R[v, 1:]].max() for v in range(1, V + 1)])
    assert (orc.min_label(g) == want_min).all()
    assert (orc.max_color(g) == want_max).all()
    # FW(p) & BW(p) = SCC(p) for every p
    for p in range(1, V + 1):
        fw = orc.reach(g, p, 0).astype(bool)
        bw = orc.reach(g, p, 1).astype(bool)
        assert (fw == R[p]).all()
        assert (bw == R[:, p]).all()
        assert ((fw & bw) == (lab == lab[p])).all()
    # trim-1 removes only singleton SCCs
    rem = orc.trim1(g).astype(bool)
    s = np.bincount(lab, minlength=V + 1)
    assert (s[lab[rem]] == 1).all()


def test_trim1_removes_every_dag_vertex():
    rng = np.random.default_rng(7)
    V = 500
    u = rng.integers(1, V, 3000)
    v = u + rng.integers(1, 50, 3000)
    keep = v <= V
    g = orc.build_csr(V, u[keep], v[keep])
    rem = orc.trim1(g)
    assert rem[1:].all()


def test_trim2_snapshot_pairs_are_two_sccs(known, fixture_path):
    g = _load(known, fixture_path, "sample128")
    lab, _ = orc.tarjan(g)
    dead = np.zeros(g.V + 1, np.uint8)
    pair = orc.trim2_snapshot(g, dead)
    hit = np.flatnonzero(pair)
    assert hit.size == 2          # sample128 has exactly one 2-vertex SCC
    a, b = hit
    assert pair[a] == b and pair[b] == a and lab[a] == lab[b]
    assert (lab == lab[a]).sum() == 2


def test_first_pivot_rule():
    # degrees: v1 out 2 in 1, v2 out 1 in 2, v3 out 1 in 1
    g = orc.build_csr(3, [1, 1, 2, 3], [2, 3, 1, 2])
    # products: v1 = 2*1 = 2, v2 = 1*2 = 2, v3 = 1*1 = 1 -> tie to smaller id
    assert orc.first_pivot(g) == 1
    dead = np.array([0, 1, 0, 0], np.uint8)
    assert orc.first_pivot(g, dead) == 2


def test_planted_cycles_are_sccs():
    V, L = 200, 37
    cyc = np.arange(10, 10 + L)
    g = orc.build_csr(V, cyc, np.roll(cyc, -1))
    lab, n = orc.tarjan(g)
    assert n == V - L + 1
    assert (lab[cyc] == 10).all()


@pytest.mark.parametrize("name", ["sample0", "sample4", "sample5", "sample1", "sample128", "sample2"])
def test_cpu_ports_equal_tarjan(known, fixture_path, name):
    g = _load(known, fixture_path, name)
    want, n = orc.tarjan(g)
    for t1, t2 in itertools.product([0, 1], [0, 1]):
        lab, m, _ = orc.vhong(g, t1, t2, threads=4)
        assert m == n and (lab == want).all(), ("vhong", t1, t2)
    for t1 in [0, 1]:
        lab, m, _ = orc.vslota(g, t1, threads=4)
        assert m == n and (lab == want).all(), ("vslota", t1)


@pytest.mark.parametrize("seed", range(10))
def test_cpu_ports_random(seed):
    rng = np.random.default_rng(1000 + seed)
    V = int(rng.integers(50, 3000))
    E = int(rng.integers(V // 2, 4 * V))
    g = orc.build_csr(V, rng.integers(1, V + 1, E), rng.integers(1, V + 1, E))
    want, n = orc.tarjan(g)
    lab, m, _ = orc.vhong(g, True, True, threads=3)
    assert m == n and (lab == want).all()
    lab, m, _ = orc.vslota(g, True, threads=3)
    assert m == n and (lab == want).all()


@pytest.mark.parametrize("idx", range(6))
def test_local_scc_known_answers(known, fixture_path, idx):
    pa = known["partitions"][idx]
    g = _load(known, fixture_path, pa["graph"])
    part = orc.load_partition(fixture_path(pa["part"]), g.V)
    assert int(part[1:g.V + 1].max()) + 1 <= pa["K"]
    src = np.repeat(np.arange(g.V + 1), np.diff(g.Fr.astype(np.int64)))
    assert int((part[src] != part[g.Fc]).sum()) == pa["cross_arcs"]
    lab, n = orc.local_scc(g, part)
    assert n == pa["local_sccs"]
    glab, gn = orc.tarjan(g)
    assert gn == pa["global_sccs"]
    pv = part[1:g.V + 1]
    spanning = sum(1 for r in np.unique(glab[1:]) if np.unique(pv[glab[1:] == r]).size > 1)
    assert spanning == pa["spanning"]
    # every local SCC lies inside one global SCC
    assert all(np.unique(glab[1:][lab[1:] == r]).size == 1 for r in np.unique(lab[1:]))
