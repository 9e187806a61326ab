"""Host-side checks that need no GPU: the C-ABI library loads and exports
every declared symbol, the GTgraph loader and CSR builder reproduce the
reference layout (checked against the oracle's restatement of load.cpp),
generators are seeded and exact, and the CLI's error paths match the
reference (help text and exit code 1)."""
import os
import re
import subprocess

import numpy as np
import pytest

import detectingscc_amd as scc
from oracle import pyoracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURES = ["sample0", "sample4", "sample5", "sample1", "sample128", "sample2"]


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "scc_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = "\n".join(ln for ln in text.splitlines() if not ln.lstrip().startswith("#"))
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "while", "for", "sizeof", "return", "defined"}
    return sorted({n for n in names if n not in skip and not n.startswith("SCC_")})


def test_library_exports_every_declared_symbol():
    L = scc.lib()
    names = _declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_status_strings_and_version():
    L = scc.lib()
    assert L.scc_status_string(0) == b"ok"
    assert b"gfx950" in L.scc_version()


@pytest.mark.parametrize("name", FIXTURES)
def test_loader_matches_oracle_layout(known, fixture_path, name):
    path = fixture_path(known["graphs"][name]["file"])
    g = scc.load_full_graph(path)
    o = orc.load_gtgraph(path)
    assert (g.V, g.E) == (o.V, o.E)
    for a in ["Fr", "Fc", "Br", "Bc"]:
        assert np.array_equal(getattr(g, a), getattr(o, a)), a
    # closed form of the layout (load.cpp:47-113)
    assert g.Fr[0] == 0 and g.Fr[1] == 0 and g.Fr[-1] == g.E and g.Br[-1] == g.E
    src = np.repeat(np.arange(g.V + 1), np.diff(g.Fr.astype(np.int64)))
    key = src.astype(np.int64) * (g.V + 1) + g.Fc
    assert (np.diff(key) >= 0).all()                      # sorted by (src, dst)
    dst = np.repeat(np.arange(g.V + 1), np.diff(g.Br.astype(np.int64)))
    key = dst.astype(np.int64) * (g.V + 1) + g.Bc
    assert (np.diff(key) >= 0).all()                      # in-lists ascending by source
    assert np.array_equal(np.sort(src.astype(np.int64) * (g.V + 1) + g.Fc),
                          np.sort(g.Bc.astype(np.int64) * (g.V + 1) + dst))


def test_loader_keeps_duplicate_arcs(known, fixture_path):
    g = scc.load_full_graph(fixture_path("sample0.gr"))
    # data/sample0.gr lists "a 4 1 1" twice
    assert list(g.Fc[g.Fr[4]:g.Fr[5]]) == [1, 1]


def _write(tmp_path, text, name="g.gr"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


HEADER = "c\nc\nc\nc\nc\nc\nc\n"


@pytest.mark.parametrize("body,why", [
    ("", "missing problem line"),
    ("p sp 3 2\na 1 2 1\n", "short arc list"),
    ("p sp 3 1\na 1 4 1\n", "id above V"),
    ("p sp 3 1\na 0 2 1\n", "id zero"),
    ("p sp 3 1\na 1 x 1\n", "non-numeric id"),
    ("p sp x 1\n", "bad V"),
])
def test_loader_rejects_malformed(tmp_path, body, why):
    with pytest.raises(scc.SCCError) as e:
        scc.load_full_graph(_write(tmp_path, HEADER + body))
    assert e.value.status == 3, why


def test_loader_missing_file():
    with pytest.raises(scc.SCCError) as e:
        scc.load_full_graph("/nonexistent/graph.gr")
    assert e.value.status == 2


def test_loader_empty_graph(tmp_path):
    g = scc.load_full_graph(_write(tmp_path, HEADER + "p sp 4 0\n"))
    assert g.V == 4 and g.E == 0 and (g.Fr == 0).all() and (g.Br == 0).all()


def test_gtgraph_roundtrip(tmp_path, known, fixture_path):
    g = scc.load_full_graph(fixture_path("sample128.gr"))
    p = str(tmp_path / "rt.gr")
    scc.write_gtgraph(p, g)
    h = scc.load_full_graph(p)
    for a in ["Fr", "Fc", "Br", "Bc"]:
        assert np.array_equal(getattr(g, a), getattr(h, a))


def test_metis_writer_format(tmp_path, fixture_path):
    g = scc.load_full_graph(fixture_path("sample0.gr"))
    p = str(tmp_path / "m.graph")
    scc.write_metis(p, g)
    lines = open(p).read().split("\n")
    assert lines[0] == "6 5"
    # vertex 1: out {2}, in {4,4}; vertex 4: out {1,1}, in {2}
    assert lines[1] == "2 4 4 "
    assert lines[4] == "1 1 2 "
    assert lines[3] == ""   # vertex 3 is isolated


@pytest.mark.parametrize("item", range(3))
def test_partition_parse_only_files(known, fixture_path, item):
    pp = known["partition_parse_only"][item]
    part = scc.load_partition(fixture_path(pp["part"]), pp["lines"])
    assert part.shape == (pp["lines"] + 2,)
    o = orc.load_partition(fixture_path(pp["part"]), pp["lines"])
    assert np.array_equal(part, o)


def test_partition_too_short(fixture_path):
    with pytest.raises(scc.SCCError):
        scc.load_partition(fixture_path("MeTiSInput0.graph.part.2"), 7)


def test_csr_from_arcs_matches_oracle():
    rng = np.random.default_rng(3)
    V, E = 3000, 20000
    s = rng.integers(1, V + 1, E)
    d = rng.integers(1, V + 1, E)
    g = scc.from_arcs(V, s, d)
    o = orc.build_csr(V, s, d)
    for a in ["Fr", "Fc", "Br", "Bc"]:
        assert np.array_equal(getattr(g, a), getattr(o, a)), a


def test_csr_rejects_bad_ids():
    with pytest.raises(scc.SCCError):
        scc.from_arcs(5, [1, 6], [2, 3])


def test_generators_seeded_and_thread_independent(monkeypatch):
    a = scc.gen_rmat(12, seed=5)
    monkeypatch.setenv("SCC_HOST_THREADS", "3")
    b = scc.gen_rmat(12, seed=5)
    monkeypatch.setenv("SCC_HOST_THREADS", "1")
    c = scc.gen_rmat(12, seed=5)
    for x in (b, c):
        assert np.array_equal(a.Fc, x.Fc) and np.array_equal(a.Fr, x.Fr)
    d = scc.gen_rmat(12, seed=6)
    assert not np.array_equal(a.Fc, d.Fc)


def test_generator_sizes():
    r = scc.gen_rmat(10)
    assert r.V == 1024 and r.E == 16 * 1024
    e = scc.gen_er(1000, 5000, cycles=3, cmin=4, cmax=4, seed=2)
    assert e.V == 1000 and 5000 - 20 < e.E - 12 <= 5000
    src = np.repeat(np.arange(e.V + 1), np.diff(e.Fr.astype(np.int64)))
    loops = int((src == e.Fc).sum())
    assert loops == 0
    gr = scc.gen_grid(64, 0.7, seed=1)
    assert gr.V == 4096
    expect = 0.7 * 4 * 64 * 63
    assert abs(gr.E - expect) < 5 * np.sqrt(expect)
    src = np.repeat(np.arange(gr.V + 1), np.diff(gr.Fr.astype(np.int64))).astype(np.int64)
    diff = np.abs(src - gr.Fc.astype(np.int64))
    assert set(np.unique(diff)) <= {1, 64}


def test_grid_full_is_one_scc():
    g = scc.gen_grid(20, 1.0, seed=1)
    lab, n = orc.tarjan(g)
    assert n == 1 and g.E == 4 * 20 * 19


def test_communities_partition():
    g = scc.gen_communities(10000, 10, deg=8.0, inter=0.02, seed=5)
    part = g.part[1:g.V + 1]
    counts = np.bincount(part)
    assert counts.size == 10 and counts.min() == counts.max() == 1000
    src = np.repeat(np.arange(g.V + 1), np.diff(g.Fr.astype(np.int64)))
    cross = (g.part[src] != g.part[g.Fc]).mean()
    assert 0.01 < cross < 0.03


def test_size_histogram():
    lab = np.array([0, 1, 1, 3, 1, 5, 5], np.uint32)
    assert scc.size_histogram(lab) == [(1, 1), (2, 1), (3, 1)]
    with pytest.raises(scc.SCCError):
        scc.size_histogram(np.array([0, 2, 2], np.uint32))    # not canonical


def _cli(*args):
    return subprocess.run([scc.CLI_PATH, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                          text=True, timeout=60)


def test_cli_help_on_too_few_args():
    r = _cli("-a", "g")
    assert r.returncode == 1 and r.stdout.startswith("To execute:")


def test_cli_unknown_algorithm(fixture_path):
    r = _cli("-a", "z", "-p", "1", "-q", "1", "-w", "1", "-f", fixture_path("sample0.gr"))
    assert r.returncode == 1 and "To execute:" in r.stdout


def test_cli_missing_file():
    r = _cli("-a", "g", "-p", "1", "-q", "1", "-w", "1", "-f", "/nonexistent.gr")
    assert r.returncode == 1 and "Error Reading graph file" in r.stdout


def test_cli_detect_needs_partition(fixture_path):
    r = _cli("-a", "d", "-p", "1", "-q", "1", "-w", "1", "-f", fixture_path("sample0.gr"))
    assert r.returncode == 1 and "Error Reading metis graph file" in r.stdout


def _check_batch_layout(b):
    assert b.row_off[0] == 0 and b.col_off[0] == 0
    for i in range(b.num_graphs):
        g = b.graph(i)
        assert g.Fr[0] == 0 and g.Fr[1] == 0 and g.Fr[-1] == g.E == b.E[i]
        assert np.all(np.diff(g.Fr.astype(np.int64)) >= 0) and np.all(np.diff(g.Br.astype(np.int64)) >= 0)
        if g.E:
            assert g.Fc.min() >= 1 and g.Fc.max() <= g.V and g.Bc.min() >= 1 and g.Bc.max() <= g.V
        # the reverse CSR holds exactly the forward arcs
        src = np.repeat(np.arange(g.V + 1), np.diff(g.Fr.astype(np.int64))[:g.V + 1])
        dst = np.repeat(np.arange(g.V + 1), np.diff(g.Br.astype(np.int64))[:g.V + 1])
        fw = sorted(zip(src.tolist(), g.Fc.tolist()))
        bw = sorted(zip(g.Bc.tolist(), dst.tolist()))
        assert fw == bw


def test_batch_mix_layout_and_seed():
    b = scc.gen_batch_mix(40, 6, 10, 2.0, 12.0, seed=9)
    again = scc.gen_batch_mix(40, 6, 10, 2.0, 12.0, seed=9)
    assert np.array_equal(b.Fc, again.Fc) and np.array_equal(b.V, again.V)
    assert b.V.min() >= 2 ** 6 - 1 and b.V.max() <= 2 ** 10 + 1
    assert set(np.unique(b.kind)) <= {0, 1, 2}
    _check_batch_layout(b)
    # per-graph arrays match a rebuild from the same arcs
    g = b.graph(7)
    src = np.repeat(np.arange(g.V + 1, dtype=np.uint32), np.diff(g.Fr.astype(np.int64))[:g.V + 1])
    h = scc.from_arcs(g.V, src, g.Fc)
    assert np.array_equal(h.Fr, g.Fr) and np.array_equal(h.Fc, g.Fc) and np.array_equal(h.Bc, g.Bc)


def test_serving_mix_distribution():
    b = scc.gen_serving_mix(300, median_v=4096, sigma=1.0, vcap=1 << 15, degree=6.0, seed=2)
    assert b.V.max() <= 1 << 15 and b.V.min() >= 2
    lg = np.log(b.V.astype(np.float64))
    assert abs(np.median(lg) - np.log(4096)) < 0.3          # lognormal median
    assert 0.7 < np.std(lg[b.V < 1 << 15]) < 1.2            # sigma, cap trimmed
    assert abs(float(b.E.sum()) / float(b.V.sum()) - 6.0) < 0.2
    assert 100 < int((b.kind == 1).sum()) < 200             # ER / R-MAT even odds
    _check_batch_layout(scc.gen_serving_mix(12, median_v=300, sigma=1.5, vcap=5000, degree=8.0, seed=3))


def test_batch_from_graphs_roundtrip():
    gs = [scc.gen_er(50 + 7 * i, 200, cycles=2, seed=i) for i in range(6)] + [scc.from_arcs(1, [], [])]
    b = scc.batch_from_graphs(gs)
    for i, g in enumerate(gs):
        h = b.graph(i)
        for name in ("Fr", "Fc", "Br", "Bc"):
            assert np.array_equal(getattr(h, name), getattr(g, name))


def test_csr_cache_roundtrip(tmp_path, fixture_path):
    for g in (scc.load_full_graph(fixture_path("sample128.gr")), scc.gen_rmat(10, seed=4),
              scc.from_arcs(3, [], [])):
        p = str(tmp_path / "g.csr")
        scc.write_csr(p, g)
        h = scc.load_graph(p)
        assert (h.V, h.E) == (g.V, g.E)
        for a in ["Fr", "Fc", "Br", "Bc"]:
            assert np.array_equal(getattr(g, a), getattr(h, a))
    # the text format still loads through the same entry point
    t = scc.load_graph(fixture_path("sample1.gr"))
    assert np.array_equal(t.Fc, scc.load_full_graph(fixture_path("sample1.gr")).Fc)


def test_csr_cache_rejects_damage(tmp_path, fixture_path):
    g = scc.load_full_graph(fixture_path("sample128.gr"))
    p = tmp_path / "g.csr"
    scc.write_csr(str(p), g)
    raw = bytearray(p.read_bytes())
    hdr = 24
    cases = {
        "truncated": bytes(raw[:-4]),
        "bad magic": b"X" + bytes(raw[1:]),
        # an out-of-range column id in Fc
        "bad id": bytes(raw[:hdr + 4 * (g.V + 2)]) + (g.V + 1).to_bytes(4, "little") + bytes(raw[hdr + 4 * (g.V + 3):]),
        # reverse CSR no longer matches: move one in-arc between vertices 1 and 2
        "reverse": None,
    }
    br_off = hdr + 4 * (g.V + 2) + 4 * g.E
    br = np.frombuffer(bytes(raw[br_off:br_off + 4 * (g.V + 2)]), dtype=np.uint32).copy()
    if br[2] > br[1]:
        br[2] -= 1
    else:
        br[2] += 1
    cases["reverse"] = bytes(raw[:br_off]) + br.tobytes() + bytes(raw[br_off + 4 * (g.V + 2):])
    for why, data in cases.items():
        q = tmp_path / f"{why.replace(' ', '_')}.csr"
        q.write_bytes(data)
        with pytest.raises(scc.SCCError) as e:
            scc.load_graph(str(q))   # without the magic it is parsed as GTgraph text
        assert e.value.status == 3, why


@pytest.mark.parametrize("n", [1, 2, 3, 5, 16, 17, 1000, 65537, 1 << 20])
def test_seeded_permutation_is_a_bijection(n):
    p = scc.seeded_permutation(n, 77)
    assert np.array_equal(np.sort(p), np.arange(n, dtype=np.uint32))
    assert np.array_equal(scc.seeded_permutation(n, 77), p)          # a pure function of (n, seed)
    if n >= 1000:   # and a real shuffle: few fixed points, a different seed differs
        assert (p == np.arange(n)).sum() < 10
        assert (scc.seeded_permutation(n, 78) != p).mean() > 0.99


def test_rmat_permutation_relabels_only():
    """The permuted R-MAT graph is the unpermuted draw relabelled: same
    degree sequence as a multiset, same arc count."""
    g = scc.gen_rmat(12, seed=3)
    assert g.E == 16 << 12
    perm = scc.seeded_permutation(1 << 12, 3)
    outdeg = np.diff(g.Fr.astype(np.int64))[1:]
    # vertex perm[r] + 1 of the permuted graph is vertex r + 1 of the draw:
    # relabelling back gives a degree sequence whose sum is E
    back = outdeg[perm]
    assert back.sum() == g.E and sorted(back.tolist()) == sorted(outdeg.tolist())
