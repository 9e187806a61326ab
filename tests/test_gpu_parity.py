"""GPU parity: HIP results through the C ABI against the pinned CPU oracle.

Every comparison is exact (integer work): canonical labels, SCC count and
size histogram equal host Tarjan; phase outputs equal their closed forms.
"""
import itertools

import numpy as np
import pytest

import detectingscc_amd as scc
from oracle import pyoracle as orc
from helpers import assert_same_labels, digest, random_graph

pytestmark = pytest.mark.gpu

FIXTURES = ["sample0", "sample4", "sample5", "sample1", "sample128", "sample2"]
ALGOS = ["g", "h", "x", "y"]


@pytest.fixture(scope="module")
def ctx():
    c = scc.Context(0)
    yield c
    c.close()


def _solve(ctx, g, **kw):
    ctx.upload(g)
    st = ctx.run(**kw)
    return ctx.labels(), st


def ref_mapping(algo):
    """The reference's own frontier mapping: a W-lane virtual warp for -a x
    and -a y, a vertex per lane otherwise (the automatic expansion is the
    direction-optimised search)."""
    return scc.EXPAND_VWARP if algo in "xy" else scc.EXPAND_VERTEX


def _check(lab, st, want, n, what):
    assert_same_labels(lab, want, what)
    assert st.num_sccs == n, what
    s = np.bincount(want[1:].astype(np.int64))
    assert st.largest == s.max(), what
    assert scc.size_histogram(lab) == scc.size_histogram(want), what


def _check_unsized(lab, st, want, n, what):
    # without sizes the Hong drivers leave the first round's labels unwritten
    # until the end (from the search's bitmaps, or a fix-up before the later
    # phases): the labels and the count must still be Tarjan's
    assert_same_labels(lab, want, what)
    assert st.num_sccs == n, what


@pytest.mark.parametrize("name", FIXTURES)
def test_fixtures_all_algorithms(ctx, known, fixture_path, name):
    g = scc.load_full_graph(fixture_path(known["graphs"][name]["file"]))
    want, n = orc.tarjan(g)
    assert digest(want) == known["graphs"][name]["digest"]
    for algo, t1, t2 in itertools.product(ALGOS, [0, 1], [0, 1]):
        for W in ([1] if algo in "gh" else [1, 2, 4, 8, 16, 32, 64]):
            lab, st = _solve(ctx, g, algo=algo, trim1=t1, trim2=t2, warp=W, expand=ref_mapping(algo))
            _check(lab, st, want, n, f"{name} -a {algo} -p {t1} -q {t2} -w {W}")
            if algo in "gx":
                lab, st = _solve(ctx, g, algo=algo, trim1=t1, trim2=t2, warp=W, sizes=False,
                                 expand=ref_mapping(algo))
                _check_unsized(lab, st, want, n, f"{name} -a {algo} -p {t1} -q {t2} -w {W} unsized")
        lab, st = _solve(ctx, g, algo=algo, trim1=t1, trim2=t2)
        _check(lab, st, want, n, f"{name} -a {algo} -p {t1} -q {t2} default expansion")


@pytest.mark.parametrize("seed", range(12))
def test_random_graphs(ctx, seed):
    rng = np.random.default_rng(seed)
    V = int(rng.integers(1, 4000))
    E = int(rng.integers(0, 6 * V + 1))
    g = random_graph(V, E, rng)
    want, n = orc.tarjan(g)
    for algo in ALGOS:
        lab, st = _solve(ctx, g, algo=algo, warp=int(rng.choice(scc.WARP_SIZES)), expand=ref_mapping(algo))
        _check(lab, st, want, n, f"seed {seed} V={V} E={E} -a {algo}")
        lab, st = _solve(ctx, g, algo=algo, expand=scc.EXPAND_LB)
        _check(lab, st, want, n, f"seed {seed} V={V} E={E} -a {algo} LB")
        lab, st = _solve(ctx, g, algo=algo, expand=scc.EXPAND_DO)
        _check(lab, st, want, n, f"seed {seed} V={V} E={E} -a {algo} DO")
        lab, st = _solve(ctx, g, algo=algo, expand=scc.EXPAND_DO, sizes=False)
        _check_unsized(lab, st, want, n, f"seed {seed} V={V} E={E} -a {algo} DO unsized")


@pytest.mark.parametrize("scale", [12, 16])
def test_rmat_lb_hubs(ctx, scale):
    """R-MAT hubs exceed the hub threshold, so the chunked hub path runs."""
    g = scc.gen_rmat(scale, seed=scale)
    assert np.diff(g.Fr.astype(np.int64)).max() >= 512
    want, n = orc.tarjan(g)
    for algo in ALGOS:
        for ex in (scc.EXPAND_LB, scc.EXPAND_DO):
            lab, st = _solve(ctx, g, algo=algo, expand=ex)
            _check(lab, st, want, n, f"rmat{scale} -a {algo} expand {ex}")


def test_reach_direction_optimised_switches(ctx):
    """A dense ER graph from a hub: the DO schedule takes bottom-up passes
    and still returns exactly the BFS reach sets, with and without a dead set."""
    g = scc.gen_er(20000, 400000, seed=5)
    ctx.upload(g)
    src = int(np.argmax(np.diff(g.Fr.astype(np.int64))[1:])) + 1
    dead = (np.arange(g.V + 1) % 11 == 4).astype(np.uint8)
    dead[src] = 0
    for d in (0, 1):
        for dd in (None, dead):
            got, levels, _ = ctx.phase_reach(src, d, scc.EXPAND_DO, 1, dead=dd)
            assert np.array_equal(got, orc.reach(g, src, d, dead=dd)), (d, dd is None)


def test_sizes_flag(ctx):
    g = scc.gen_rmat(12, seed=2)
    ctx.upload(g)
    want, n = orc.tarjan(g)
    big = int(np.bincount(want[1:].astype(np.int64)).max())
    assert ctx.run(algo="g", sizes=True).largest == big
    st = ctx.run(algo="g", sizes=False)
    assert st.largest == 0 and st.num_sccs == n


@pytest.mark.parametrize("expand,warp", [(scc.EXPAND_VERTEX, 1), (scc.EXPAND_LB, 1),
                                         (scc.EXPAND_DO, 1)] +
                         [(scc.EXPAND_VWARP, w) for w in scc.WARP_SIZES])
def test_reach_matches_bfs(ctx, expand, warp):
    g = scc.gen_rmat(14, seed=3)
    ctx.upload(g)
    rng = np.random.default_rng(warp)
    dead = (rng.random(g.V + 1) < 0.1).astype(np.uint8)
    dead[0] = 0
    src = int(np.flatnonzero(dead[1:] == 0)[np.argmax(np.diff(g.Fr.astype(np.int64))[1:][dead[1:] == 0])]) + 1
    for d in (0, 1):
        got, levels, _ = ctx.phase_reach(src, d, expand, warp, dead=dead)
        want = orc.reach(g, src, d, dead=dead)
        assert np.array_equal(got, want), (expand, warp, d)
        assert levels >= 1


def test_trim1_fixpoint(ctx, fixture_path):
    for g in [scc.load_full_graph(fixture_path("sample128.gr")), scc.gen_rmat(13, seed=9),
              scc.gen_er(5000, 8000, seed=4)]:
        ctx.upload(g)
        assert np.array_equal(ctx.phase_trim1(), orc.trim1(g))
        dead = (np.arange(g.V + 1) % 7 == 3).astype(np.uint8)
        assert np.array_equal(ctx.phase_trim1(dead), orc.trim1(g, dead))


def test_trim2_pairs(ctx, fixture_path):
    for g in [scc.load_full_graph(fixture_path("sample128.gr")), scc.gen_er(3000, 3500, seed=8)]:
        ctx.upload(g)
        want_lab, _ = orc.tarjan(g)
        dead = orc.trim1(g)
        got = ctx.phase_trim2(dead)
        snap = orc.trim2_snapshot(g, dead)
        hit = np.flatnonzero(got)
        sizes = np.bincount(want_lab.astype(np.int64), minlength=g.V + 1)
        # every removed vertex is in a Tarjan SCC of size <= 2 whose min is its label
        assert (sizes[want_lab[hit]] <= 2).all()
        assert (got[hit] == want_lab[hit]).all()
        # every pair found on the snapshot is found by the concurrent pass
        assert (got[np.flatnonzero(snap)] != 0).all()


def test_first_pivot(ctx):
    for seed in range(3):
        g = scc.gen_rmat(12, seed=seed)
        ctx.upload(g)
        assert ctx.phase_first_pivot() == orc.first_pivot(g)
        dead = (np.arange(g.V + 1) % 3 == 0).astype(np.uint8)
        assert ctx.phase_first_pivot(dead) == orc.first_pivot(g, dead)


def test_min_label_and_color(ctx):
    for g in [scc.gen_rmat(12, seed=1), scc.gen_grid(40, 0.7, seed=2), scc.gen_er(3000, 6000, seed=3)]:
        ctx.upload(g)
        dead = (np.arange(g.V + 1) % 5 == 1).astype(np.uint8)
        for d in (None, dead):
            assert np.array_equal(ctx.phase_min_label(d), orc.min_label(g, d))
            assert np.array_equal(ctx.phase_max_color(d), orc.max_color(g, d))


def test_config1_er_planted(ctx):
    """ER 16,384 V / 131,072 draws + 64 planted cycles of length 2..256."""
    g = scc.gen_er(16384, 131072, cycles=64, cmin=2, cmax=256, seed=1)
    want, n = orc.tarjan(g)
    for algo, t1, t2 in itertools.product(ALGOS, [0, 1], [0, 1]):
        lab, st = _solve(ctx, g, algo=algo, trim1=t1, trim2=t2, warp=8)
        _check(lab, st, want, n, f"C1 -a {algo} -p {t1} -q {t2}")


def test_grid_high_diameter(ctx):
    g = scc.gen_grid(256, 0.7, seed=4096)
    want, n = orc.tarjan(g)
    for algo in ALGOS:
        lab, st = _solve(ctx, g, algo=algo, warp=4)
        _check(lab, st, want, n, f"grid256 -a {algo}")


def test_repeatable(ctx):
    g = scc.gen_rmat(15, seed=11)
    ctx.upload(g)
    first = None
    for algo in ALGOS:
        for _ in range(2):
            ctx.run(algo=algo, warp=16)
            lab = ctx.labels()
            if first is None:
                first = lab
            assert np.array_equal(lab, first)


def test_dropin_wrappers_print_reference_lines(ctx, capfd, fixture_path):
    g = scc.load_full_graph(fixture_path("sample128.gr"))
    rc = scc.vHong(g.E, g.V, g.Fc, g.Fr, g.Bc, g.Br, True, True)
    out = capfd.readouterr().out
    assert rc == 0
    lines = out.split("\n")
    assert lines[0].startswith("numberOf1Sccs: ")
    assert lines[1] == "" and lines[2] == "Number Of Sccs : 13" and lines[3].startswith("Time : ")
    rc = scc.wSlota(g.E, g.V, g.Fc, g.Fr, g.Bc, g.Br, True, False, 64)
    out = capfd.readouterr().out
    assert rc == 0 and out.split("\n")[1] == "Number Of Sccs : 13"


@pytest.mark.parametrize("seed", range(6))
def test_device_histogram(ctx, seed):
    """The size histogram computed on the device equals the one counted on
    the host from Tarjan's labels, for every algorithm, with and without the
    sizes pass of the run itself."""
    rng = np.random.default_rng(900 + seed)
    V = int(rng.integers(1, 20000))
    g = random_graph(V, int(rng.integers(0, 3 * V + 1)), rng)
    want, n = orc.tarjan(g)
    sizes = np.bincount(want[1:].astype(np.int64))
    hist = np.bincount(sizes[sizes > 0])
    pairs = [(k, int(c)) for k, c in enumerate(hist) if c]
    ctx.upload(g)
    for algo, sized in [("g", True), ("g", False), ("h", True), ("x", False), ("y", True)]:
        ctx.run(algo=algo, warp=8, sizes=sized)
        got, ms = ctx.histogram()
        assert got == pairs, f"seed {seed} -a {algo} sizes={sized}"
        assert ms >= 0
