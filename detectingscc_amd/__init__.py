"""MI355X-native strongly-connected-components engine (Python host API).

Thin ctypes binding of ``libscc_amd.so`` (C ABI in ``include/scc_amd.h``).
Every SCC computation runs in the HIP library on the GPU; this module only
moves numpy arrays across the boundary.  Importing the package does not
touch the GPU; creating a :class:`Context` does, and fails loudly when the
library or a device is missing.

Names mirror the reference (Amit104/DetectingSCC):
  - :func:`load_full_graph`   <- ``loadFullGraph`` (src/load.cpp:12-140)
  - :func:`vHong`, :func:`wHong`, :func:`vSlota`, :func:`wSlota`,
    :func:`detectSCC`         <- the drivers of include/scc.h:49-53
Graphs use the reference layout: vertices 1..V, dummy vertex 0, row arrays of
V+2 entries, column arrays of E entries.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libscc_amd.so")
CLI_PATH = os.path.join(_PKG, "bin", "scc")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u64p = ctypes.POINTER(ctypes.c_uint64)

EXPAND_AUTO, EXPAND_VERTEX, EXPAND_VWARP, EXPAND_LB, EXPAND_DO = 0, 1, 2, 3, 4
PHASES = ["trim1", "trim2", "pivot", "reach", "update", "wcc", "color", "local", "merge", "label"]
WARP_SIZES = (1, 2, 4, 8, 16, 32, 64)


class SCCError(RuntimeError):
    """A non-zero scc_status from the library."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


class _HostGraph(ctypes.Structure):
    _fields_ = [("V", ctypes.c_uint32), ("E", ctypes.c_uint32),
                ("Fr", _u32p), ("Fc", _u32p), ("Br", _u32p), ("Bc", _u32p)]


class _Opts(ctypes.Structure):
    _fields_ = [("algo", ctypes.c_char), ("trim1", ctypes.c_int), ("trim2", ctypes.c_int),
                ("warp", ctypes.c_int), ("expand", ctypes.c_int), ("block", ctypes.c_int),
                ("check_every", ctypes.c_int), ("sizes", ctypes.c_int), ("phase_times", ctypes.c_int)]


class _Stats(ctypes.Structure):
    _fields_ = [("num_sccs", ctypes.c_uint32), ("num_trim1", ctypes.c_uint32),
                ("num_trim2", ctypes.c_uint32), ("num_pivot", ctypes.c_uint32),
                ("largest", ctypes.c_uint32), ("rounds", ctypes.c_uint32),
                ("levels", ctypes.c_uint32), ("first_pivot", ctypes.c_uint32),
                ("first_fw_vertices", ctypes.c_uint64), ("first_fw_arcs", ctypes.c_uint64),
                ("first_bw_vertices", ctypes.c_uint64), ("first_bw_arcs", ctypes.c_uint64),
                ("ms_scc", ctypes.c_double), ("ms_first_reach", ctypes.c_double),
                ("ms_phase", ctypes.c_double * len(PHASES)), ("first_work", ctypes.c_uint64 * 8)]


class _BatchHost(ctypes.Structure):
    _fields_ = [("num_graphs", ctypes.c_uint32), ("V", _u32p), ("E", _u32p),
                ("row_off", _u64p), ("col_off", _u64p),
                ("Fr", _u32p), ("Fc", _u32p), ("Br", _u32p), ("Bc", _u32p), ("kind", _u8p)]


class _BatchOpts(ctypes.Structure):
    _fields_ = [("algo", ctypes.c_char), ("trim1", ctypes.c_int), ("trim2", ctypes.c_int),
                ("lanes", ctypes.c_int), ("block", ctypes.c_int), ("autotuned", ctypes.c_int)]


class _BatchStats(ctypes.Structure):
    _fields_ = [("graphs", ctypes.c_uint32), ("ms", ctypes.c_double), ("graphs_per_s", ctypes.c_double),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("vertices", ctypes.c_uint64), ("arcs", ctypes.c_uint64), ("launches", ctypes.c_uint32)]


class _ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("rejected", ctypes.c_uint64), ("failed", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("solo", ctypes.c_uint64),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double),
                ("graphs_per_s", ctypes.c_double), ("peak_bytes", ctypes.c_uint64)]


_LIB = None


def lib() -> ctypes.CDLL:
    """Load libscc_amd.so once; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m detectingscc_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name in ["scc_status_string", "scc_last_error", "scc_version"]:
            getattr(L, name).restype = ctypes.c_char_p
        L.scc_status_string.argtypes = [ctypes.c_int]
        vp = ctypes.c_void_p
        L.scc_gen_batch_mix.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(_BatchHost)]
        L.scc_batch_autotune.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
        L.scc_server_start.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_uint32, ctypes.POINTER(_BatchOpts)]
        L.scc_gen_serving_mix.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_uint32,
                                          ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(_BatchHost)]
        L.scc_server_submit.argtypes = [vp, ctypes.POINTER(_HostGraph), ctypes.POINTER(ctypes.c_uint64)]
        L.scc_server_poll.argtypes = [vp, ctypes.c_uint64, _u32p, _u32p, ctypes.POINTER(ctypes.c_double)]
        _LIB = L
    return _LIB


def _check(status: int) -> None:
    if status != 0:
        msg = lib().scc_last_error().decode(errors="replace")
        raise SCCError(status, msg or lib().scc_status_string(status).decode())


def _p32(a: np.ndarray):
    return a.ctypes.data_as(_u32p)


def _p8(a):
    return None if a is None else a.ctypes.data_as(_u8p)


@dataclass
class Graph:
    """Host CSR + reverse CSR in the reference layout (uint32 arrays)."""
    V: int
    E: int
    Fr: np.ndarray
    Fc: np.ndarray
    Br: np.ndarray
    Bc: np.ndarray
    part: np.ndarray | None = field(default=None, repr=False)

    def out_degree(self) -> np.ndarray:
        return np.diff(self.Fr).astype(np.int64)

    def in_degree(self) -> np.ndarray:
        return np.diff(self.Br).astype(np.int64)


def _take(hg: _HostGraph) -> Graph:
    V, E = int(hg.V), int(hg.E)
    def arr(p, n):
        return np.ctypeslib.as_array(p, shape=(max(n, 1),))[:n].copy() if n else np.zeros(0, np.uint32)
    g = Graph(V, E, arr(hg.Fr, V + 2), arr(hg.Fc, E), arr(hg.Br, V + 2), arr(hg.Bc, E))
    lib().scc_host_graph_free(ctypes.byref(hg))
    return g


def load_full_graph(path: str) -> Graph:
    """GTgraph text -> Graph (loadFullGraph, src/load.cpp:12-113)."""
    hg = _HostGraph()
    _check(lib().scc_load_gtgraph(os.fsencode(path), ctypes.byref(hg)))
    return _take(hg)


def load_partition(path: str, V: int) -> np.ndarray:
    """Partition vector file -> uint32[V+2] with ids at 1..V (load.cpp:117-130)."""
    part = np.zeros(V + 2, np.uint32)
    _check(lib().scc_load_partition(os.fsencode(path), ctypes.c_uint32(V), _p32(part)))
    return part


def from_arcs(V: int, src, dst) -> Graph:
    """Arc list (1-based ids) -> Graph in the reference layout."""
    s = np.ascontiguousarray(src, dtype=np.uint32)
    d = np.ascontiguousarray(dst, dtype=np.uint32)
    if s.shape != d.shape:
        raise ValueError("src and dst differ in length")
    hg = _HostGraph()
    _check(lib().scc_csr_from_arcs(ctypes.c_uint32(V), ctypes.c_uint64(s.size), _p32(s), _p32(d),
                                   ctypes.byref(hg)))
    return _take(hg)


def gen_er(V: int, draws: int, cycles: int = 0, cmin: int = 2, cmax: int = 256, seed: int = 1) -> Graph:
    hg = _HostGraph()
    _check(lib().scc_gen_er(ctypes.c_uint32(V), ctypes.c_uint64(draws), ctypes.c_uint32(cycles),
                            ctypes.c_uint32(cmin), ctypes.c_uint32(cmax), ctypes.c_uint64(seed),
                            ctypes.byref(hg)))
    return _take(hg)


def gen_rmat(scale: int, E: int | None = None, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 22) -> Graph:
    E = 16 * (1 << scale) if E is None else E
    hg = _HostGraph()
    _check(lib().scc_gen_rmat(ctypes.c_uint32(scale), ctypes.c_uint64(E), ctypes.c_double(a),
                              ctypes.c_double(b), ctypes.c_double(c), ctypes.c_uint64(seed),
                              ctypes.byref(hg)))
    return _take(hg)


def gen_grid(side: int, p: float = 0.7, seed: int = 4096) -> Graph:
    hg = _HostGraph()
    _check(lib().scc_gen_grid(ctypes.c_uint32(side), ctypes.c_double(p), ctypes.c_uint64(seed),
                              ctypes.byref(hg)))
    return _take(hg)


def gen_communities(V: int, K: int, deg: float = 12.0, inter: float = 0.02, seed: int = 55) -> Graph:
    hg = _HostGraph()
    part = np.zeros(V + 2, np.uint32)
    _check(lib().scc_gen_communities(ctypes.c_uint32(V), ctypes.c_uint32(K), ctypes.c_double(deg),
                                     ctypes.c_double(inter), ctypes.c_uint64(seed), ctypes.byref(hg),
                                     _p32(part)))
    g = _take(hg)
    g.part = part
    return g


def seeded_permutation(n: int, seed: int) -> np.ndarray:
    """The generators' seeded permutation of [0, n) (i -> perm[i])."""
    out = np.zeros(max(n, 1), np.uint32)
    _check(lib().scc_seeded_permutation(ctypes.c_uint32(n), ctypes.c_uint64(seed), _p32(out)))
    return out[:n]


def write_gtgraph(path: str, g: Graph) -> None:
    hg = _HostGraph(g.V, g.E, _p32(g.Fr), _p32(g.Fc), _p32(g.Br), _p32(g.Bc))
    _check(lib().scc_write_gtgraph(os.fsencode(path), ctypes.byref(hg)))


def write_csr(path: str, g: Graph) -> None:
    """Binary CSR cache: the four reference-layout arrays behind a header."""
    hg = _HostGraph(g.V, g.E, _p32(g.Fr), _p32(g.Fc), _p32(g.Br), _p32(g.Bc))
    _check(lib().scc_write_csr(os.fsencode(path), ctypes.byref(hg)))


def load_graph(path: str) -> Graph:
    """GTgraph text or binary CSR cache, told apart by the cache's magic."""
    hg = _HostGraph()
    _check(lib().scc_load_graph(os.fsencode(path), ctypes.byref(hg)))
    return _take(hg)


def write_metis(path: str, g: Graph) -> None:
    hg = _HostGraph(g.V, g.E, _p32(g.Fr), _p32(g.Fc), _p32(g.Br), _p32(g.Bc))
    _check(lib().scc_write_metis(os.fsencode(path), ctypes.byref(hg)))


def size_histogram(labels: np.ndarray) -> list[tuple[int, int]]:
    """(size, count) pairs ascending by size from canonical labels (V+1 entries)."""
    lab = np.ascontiguousarray(labels, dtype=np.uint32)
    V = lab.size - 1
    n = ctypes.c_uint32(0)
    _check(lib().scc_size_histogram(ctypes.c_uint32(V), _p32(lab), None, ctypes.c_uint32(0), ctypes.byref(n)))
    pairs = np.zeros(2 * max(n.value, 1), np.uint64)
    _check(lib().scc_size_histogram(ctypes.c_uint32(V), _p32(lab), pairs.ctypes.data_as(_u64p),
                                    ctypes.c_uint32(n.value), ctypes.byref(n)))
    return [(int(pairs[2 * i]), int(pairs[2 * i + 1])) for i in range(n.value)]


@dataclass
class Stats:
    num_sccs: int
    num_trim1: int
    num_trim2: int
    num_pivot: int
    largest: int
    rounds: int
    levels: int
    first_pivot: int
    first_fw_vertices: int
    first_fw_arcs: int
    first_bw_vertices: int
    first_bw_arcs: int
    ms_scc: float
    ms_first_reach: float
    ms_phase: dict
    first_work: tuple = ()

    # bytes of each first-search work unit (scc_stats.first_work): top-down
    # vertex 12 (two row offsets + queue entry), arc 8 (column + visited
    # word), claim 4 (queue entry); bottom-up swept vertex 4.125 (its visited
    # bit and its first-parent word: every vertex of a dense sweep, only the
    # candidates of a sparse one, whose other vertices read just their bit,
    # at most V/8 bytes per level and direction left uncounted), candidate 0
    # (its first parent is in the sweep's read), first parent probed 8
    # (bitmap pair), row after a miss 8, further parent 12
    WORK_BYTES = (12.0, 8.0, 4.0, 4.125, 0.0, 8.0, 8.0, 12.0)

    @property
    def first_work_bytes(self) -> float:
        """Algorithmic bytes of the work the first search did (counted on the
        device in a run with sizes and direction-optimised expansion)."""
        return float(sum(n * b for n, b in zip(self.first_work, self.WORK_BYTES)))


def _stats(s: _Stats) -> Stats:
    return Stats(s.num_sccs, s.num_trim1, s.num_trim2, s.num_pivot, s.largest, s.rounds, s.levels,
                 s.first_pivot, s.first_fw_vertices, s.first_fw_arcs, s.first_bw_vertices,
                 s.first_bw_arcs, s.ms_scc, s.ms_first_reach,
                 {PHASES[i]: float(s.ms_phase[i]) for i in range(len(PHASES))},
                 tuple(int(x) for x in s.first_work))


def make_opts(algo: str = "g", trim1: bool = True, trim2: bool = True, warp: int = 1,
              expand: int = EXPAND_AUTO, check_every: int = 0, sizes: bool = True,
              phase_times: bool = False) -> _Opts:
    o = _Opts()
    lib().scc_opts_default(ctypes.byref(o))
    o.algo = algo.encode()[:1]
    o.trim1, o.trim2, o.warp, o.expand, o.check_every = int(trim1), int(trim2), int(warp), int(expand), int(check_every)
    o.sizes = int(sizes)
    o.phase_times = int(phase_times)
    return o


@dataclass
class Batch:
    """Many graphs in one host layout: graph i's rows start at row_off[i]
    (V[i]+2 entries, local ids 1..V[i]), its columns at col_off[i]."""
    V: np.ndarray
    E: np.ndarray
    row_off: np.ndarray
    col_off: np.ndarray
    Fr: np.ndarray
    Fc: np.ndarray
    Br: np.ndarray
    Bc: np.ndarray
    kind: np.ndarray | None = None

    @property
    def num_graphs(self) -> int:
        return int(self.V.size)

    def graph(self, i: int) -> Graph:
        r, c = int(self.row_off[i]), int(self.col_off[i])
        V, E = int(self.V[i]), int(self.E[i])
        return Graph(V, E, self.Fr[r:r + V + 2].copy(), self.Fc[c:c + E].copy(),
                     self.Br[r:r + V + 2].copy(), self.Bc[c:c + E].copy())

    def _host(self) -> _BatchHost:
        h = _BatchHost()
        h.num_graphs = self.num_graphs
        h.V, h.E = _p32(self.V), _p32(self.E)
        h.row_off = self.row_off.ctypes.data_as(_u64p)
        h.col_off = self.col_off.ctypes.data_as(_u64p)
        h.Fr, h.Fc, h.Br, h.Bc = _p32(self.Fr), _p32(self.Fc), _p32(self.Br), _p32(self.Bc)
        h.kind = None
        return h


def batch_from_graphs(graphs: list[Graph]) -> Batch:
    V = np.array([g.V for g in graphs], np.uint32)
    E = np.array([g.E for g in graphs], np.uint32)
    ro = np.zeros(len(graphs) + 1, np.uint64)
    co = np.zeros(len(graphs) + 1, np.uint64)
    ro[1:] = np.cumsum(V.astype(np.uint64) + 2)
    co[1:] = np.cumsum(E.astype(np.uint64))
    cat = lambda xs: np.ascontiguousarray(np.concatenate(xs) if xs else np.zeros(0), dtype=np.uint32)
    return Batch(V, E, ro, co, cat([g.Fr for g in graphs]), cat([g.Fc for g in graphs]),
                 cat([g.Br for g in graphs]), cat([g.Bc for g in graphs]))


def _take_batch(h: _BatchHost) -> Batch:
    G = int(h.num_graphs)
    def arr(p, n, dt):
        return np.ctypeslib.as_array(p, shape=(max(n, 1),))[:n].astype(dt, copy=True)
    ro = arr(h.row_off, G + 1, np.uint64)
    co = arr(h.col_off, G + 1, np.uint64)
    b = Batch(arr(h.V, G, np.uint32), arr(h.E, G, np.uint32), ro, co,
              arr(h.Fr, int(ro[-1]), np.uint32), arr(h.Fc, int(co[-1]), np.uint32),
              arr(h.Br, int(ro[-1]), np.uint32), arr(h.Bc, int(co[-1]), np.uint32),
              arr(h.kind, G, np.uint8) if h.kind else None)
    lib().scc_batch_host_free(ctypes.byref(h))
    return b


def gen_batch_mix(num_graphs: int, lgmin: int = 8, lgmax: int = 16, dmin: float = 2.0,
                  dmax: float = 16.0, seed: int = 4096) -> Batch:
    """The batched mix: V log-uniform in [2^lgmin, 2^lgmax], mean degree
    uniform in [dmin, dmax], 40% ER / 40% R-MAT / 20% planted cycles + ER."""
    h = _BatchHost()
    _check(lib().scc_gen_batch_mix(num_graphs, lgmin, lgmax, dmin, dmax, seed, ctypes.byref(h)))
    return _take_batch(h)


def gen_serving_mix(num_graphs: int, median_v: float = 16384, sigma: float = 1.5, vcap: int = 1 << 22,
                    degree: float = 8.0, seed: int = 8) -> Batch:
    """Serving requests: V lognormal (median, sigma) clamped to [2, vcap],
    mean degree `degree`, half ER and half R-MAT."""
    h = _BatchHost()
    _check(lib().scc_gen_serving_mix(num_graphs, median_v, sigma, vcap, degree, seed, ctypes.byref(h)))
    return _take_batch(h)


@dataclass
class BatchStats:
    graphs: int
    ms: float
    graphs_per_s: float
    p50_ms: float
    p99_ms: float
    vertices: int
    arcs: int
    launches: int


@dataclass
class ServerStats:
    submitted: int
    completed: int
    rejected: int
    failed: int
    batches: int
    solo: int
    p50_ms: float
    p99_ms: float
    mean_ms: float
    graphs_per_s: float
    peak_bytes: int


def _batch_stats(s: _BatchStats) -> BatchStats:
    return BatchStats(s.graphs, s.ms, s.graphs_per_s, s.p50_ms, s.p99_ms, s.vertices, s.arcs, s.launches)


def make_batch_opts(algo: str = "g", trim1: bool = True, trim2: bool = True, lanes: int = 4,
                    block: int = 256, autotuned: bool = False) -> _BatchOpts:
    o = _BatchOpts()
    o.algo = algo.encode()[:1]
    o.trim1, o.trim2, o.lanes, o.block, o.autotuned = int(trim1), int(trim2), int(lanes), int(block), int(autotuned)
    return o


class Context:
    """A device context: resident graph + state on one HIP device."""

    def __init__(self, device: int = 0):
        self._ctx = ctypes.c_void_p()
        _check(lib().scc_ctx_create(ctypes.c_int(device), ctypes.byref(self._ctx)))
        self.graph: Graph | None = None
        self._batch: Batch | None = None
        self._served: dict[int, int] = {}

    def close(self) -> None:
        if self._ctx:
            lib().scc_ctx_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, g: Graph) -> None:
        _check(lib().scc_graph_upload(self._ctx, ctypes.c_uint32(g.V), ctypes.c_uint32(g.E),
                                      _p32(g.Fr), _p32(g.Fc), _p32(g.Br), _p32(g.Bc)))
        self.graph = g
        if g.part is not None:
            self.upload_partition(g.part)

    # -- generators on the device (bit-identical to gen_* on the host) --
    def _adopt(self, V: int, E: int, part=None) -> None:
        z = np.zeros(0, np.uint32)
        self.graph = Graph(V, E, z, z, z, z, part)

    def gen_er_dev(self, V: int, draws: int, cycles: int = 0, cmin: int = 2, cmax: int = 256, seed: int = 1):
        _check(lib().scc_gen_er_dev(self._ctx, ctypes.c_uint32(V), ctypes.c_uint64(draws), ctypes.c_uint32(cycles),
                                    ctypes.c_uint32(cmin), ctypes.c_uint32(cmax), ctypes.c_uint64(seed)))
        self._adopt(V, self._device_E())

    def gen_rmat_dev(self, scale: int, E: int | None = None, a: float = 0.57, b: float = 0.19, c: float = 0.19,
                     seed: int = 22):
        E = 16 * (1 << scale) if E is None else E
        _check(lib().scc_gen_rmat_dev(self._ctx, ctypes.c_uint32(scale), ctypes.c_uint64(E), ctypes.c_double(a),
                                      ctypes.c_double(b), ctypes.c_double(c), ctypes.c_uint64(seed)))
        self._adopt(1 << scale, E)

    def gen_grid_dev(self, side: int, p: float = 0.7, seed: int = 4096):
        _check(lib().scc_gen_grid_dev(self._ctx, ctypes.c_uint32(side), ctypes.c_double(p), ctypes.c_uint64(seed)))
        self._adopt(side * side, self._device_E())

    def gen_communities_dev(self, V: int, K: int, deg: float = 12.0, inter: float = 0.02, seed: int = 55):
        part = np.zeros(V + 2, np.uint32)
        _check(lib().scc_gen_communities_dev(self._ctx, ctypes.c_uint32(V), ctypes.c_uint32(K), ctypes.c_double(deg),
                                             ctypes.c_double(inter), ctypes.c_uint64(seed), _p32(part)))
        self._adopt(V, self._device_E(), part)

    def gen_batch_mix_dev(self, num_graphs: int, lgmin: int = 8, lgmax: int = 16, dmin: float = 2.0,
                          dmax: float = 16.0, seed: int = 4096, copy: bool = False):
        """The batched mix generated on the device as the context's batch;
        with copy, also returned as a host Batch (for checking)."""
        h = _BatchHost()
        _check(lib().scc_gen_batch_mix_dev(self._ctx, ctypes.c_uint32(num_graphs), ctypes.c_uint32(lgmin),
                                           ctypes.c_uint32(lgmax), ctypes.c_double(dmin), ctypes.c_double(dmax),
                                           ctypes.c_uint64(seed), ctypes.byref(h) if copy else None))
        if copy:
            self._batch = _take_batch(h)
            return self._batch
        self._batch = self.batch_info()
        return None

    def batch_info(self) -> Batch:
        """The context's batch sizes and offsets (no arrays)."""
        G = ctypes.c_uint32(0)
        _check(lib().scc_batch_info(self._ctx, ctypes.byref(G), None, None, None, None))
        n = G.value
        V = np.zeros(n, np.uint32)
        E = np.zeros(n, np.uint32)
        ro = np.zeros(n + 1, np.uint64)
        co = np.zeros(n + 1, np.uint64)
        _check(lib().scc_batch_info(self._ctx, ctypes.byref(G), _p32(V), _p32(E), ro.ctypes.data_as(_u64p),
                                    co.ctypes.data_as(_u64p)))
        z = np.zeros(0, np.uint32)
        return Batch(V, E, ro, co, z, z, z, z)

    def _device_E(self) -> int:
        V = ctypes.c_uint32(0)
        E = ctypes.c_uint32(0)
        _check(lib().scc_graph_size(self._ctx, ctypes.byref(V), ctypes.byref(E)))
        return E.value

    def download(self) -> Graph:
        """The context's device graph (CSR + reverse CSR) copied to the host."""
        hg = _HostGraph()
        _check(lib().scc_graph_download(self._ctx, ctypes.byref(hg)))
        return _take(hg)

    def upload_partition(self, part: np.ndarray) -> None:
        part = np.ascontiguousarray(part, dtype=np.uint32)
        _check(lib().scc_partition_upload(self._ctx, _p32(part)))

    def run(self, algo: str = "g", trim1: bool = True, trim2: bool = True, warp: int = 1,
            expand: int = EXPAND_AUTO, check_every: int = 0, sizes: bool = True,
            phase_times: bool = False) -> Stats:
        o = make_opts(algo, trim1, trim2, warp, expand, check_every, sizes, phase_times)
        s = _Stats()
        _check(lib().scc_run(self._ctx, ctypes.byref(o), ctypes.byref(s)))
        return _stats(s)

    def labels(self) -> np.ndarray:
        lab = np.zeros(self.graph.V + 1, np.uint32)
        _check(lib().scc_labels(self._ctx, _p32(lab)))
        return lab

    def histogram(self) -> tuple[list[tuple[int, int]], float]:
        """SCC size histogram of the last run computed on the device:
        [(size, count), ...] ascending by size, and the device ms."""
        V = self.graph.V
        cap = 2
        while cap * (cap - 1) // 2 <= V:
            cap += 1
        pairs = np.zeros(2 * cap, np.uint64)
        n = ctypes.c_uint32(0)
        ms = ctypes.c_double(0)
        _check(lib().scc_histogram(self._ctx, pairs.ctypes.data_as(_u64p), ctypes.c_uint32(cap),
                                   ctypes.byref(n), ctypes.byref(ms)))
        k = min(n.value, cap)
        return [(int(pairs[2 * i]), int(pairs[2 * i + 1])) for i in range(k)], ms.value

    def condensation(self) -> tuple[Graph, np.ndarray]:
        """The DAG of SCCs of the last run (node k = k-th smallest canonical
        label) and each vertex's node (V+1 entries)."""
        node_of = np.zeros(self.graph.V + 1, np.uint32)
        hg = _HostGraph()
        _check(lib().scc_condensation(self._ctx, ctypes.byref(hg), _p32(node_of)))
        return _take(hg), node_of

    def solve(self, g: Graph, **kw) -> tuple[np.ndarray, Stats]:
        self.upload(g)
        st = self.run(**kw)
        return self.labels(), st

    def copy_bench(self, nbytes: int, iters: int = 5) -> float:
        """Mean ms of a plain device copy of nbytes (counter calibration)."""
        ms = ctypes.c_double(0)
        _check(lib().scc_copy_bench(self._ctx, ctypes.c_uint64(nbytes), ctypes.c_int(iters),
                                    ctypes.byref(ms)))
        return ms.value

    # -- batches --
    def batch_upload(self, b: Batch) -> None:
        self._batch = b
        _check(lib().scc_batch_upload(self._ctx, ctypes.byref(b._host())))

    def batch_run(self, algo: str = "g", trim1: bool = True, trim2: bool = True, lanes: int = 4,
                  block: int = 256, autotuned: bool = False) -> BatchStats:
        o = make_batch_opts(algo, trim1, trim2, lanes, block, autotuned)
        s = _BatchStats()
        _check(lib().scc_batch_run(self._ctx, ctypes.byref(o), ctypes.byref(s)))
        return _batch_stats(s)

    def batch_run_serial(self, **kw) -> BatchStats:
        kw.setdefault("sizes", False)
        o = make_opts(**kw)
        s = _BatchStats()
        _check(lib().scc_batch_run_serial(self._ctx, ctypes.byref(o), ctypes.byref(s)))
        return _batch_stats(s)

    def batch_autotune(self, sample: int = 16) -> str:
        buf = ctypes.create_string_buffer(8192)
        _check(lib().scc_batch_autotune(self._ctx, sample, buf, len(buf)))
        return buf.value.decode()

    def batch_labels(self) -> tuple[list[np.ndarray], np.ndarray]:
        b = self._batch
        lab = np.zeros(int(b.row_off[-1]), np.uint32)
        nscc = np.zeros(b.num_graphs, np.uint32)
        _check(lib().scc_batch_labels(self._ctx, _p32(lab), _p32(nscc)))
        return [lab[int(b.row_off[i]):int(b.row_off[i]) + int(b.V[i]) + 1] for i in range(b.num_graphs)], nscc

    # -- continuous-batching server --
    def server_start(self, budget_bytes: int, max_batch: int = 256, slots: int = 2, solo_vertices: int = 0,
                     algo: str = "g", trim1: bool = True, trim2: bool = True, lanes: int = 4,
                     block: int = 256) -> None:
        o = make_batch_opts(algo, trim1, trim2, lanes, block, False)
        _check(lib().scc_server_start(self._ctx, budget_bytes, max_batch, slots, solo_vertices, ctypes.byref(o)))
        self._served = {}

    def server_submit(self, g: Graph) -> int:
        hg = _HostGraph(g.V, g.E, _p32(g.Fr), _p32(g.Fc), _p32(g.Br), _p32(g.Bc))
        rid = ctypes.c_uint64(0)
        _check(lib().scc_server_submit(self._ctx, ctypes.byref(hg), ctypes.byref(rid)))
        self._served[rid.value] = g.V
        return rid.value

    def server_poll(self, rid: int):
        """None while pending; (labels, nscc, latency_ms) once done."""
        lab = np.zeros(self._served[rid] + 1, np.uint32)
        n = ctypes.c_uint32(0)
        ms = ctypes.c_double(0)
        st = lib().scc_server_poll(self._ctx, rid, _p32(lab), ctypes.byref(n), ctypes.byref(ms))
        if st == 8:
            return None
        del self._served[rid]
        _check(st)
        return lab, n.value, ms.value

    def server_wait_all(self) -> None:
        _check(lib().scc_server_wait_all(self._ctx))

    def server_stats(self) -> ServerStats:
        s = _ServerStats()
        _check(lib().scc_server_stats_get(self._ctx, ctypes.byref(s)))
        return ServerStats(s.submitted, s.completed, s.rejected, s.failed, s.batches, s.solo, s.p50_ms, s.p99_ms,
                           s.mean_ms, s.graphs_per_s, s.peak_bytes)

    def server_stop(self) -> None:
        _check(lib().scc_server_stop(self._ctx))

    # -- single phases (kernel-level checks) --
    def _dead(self, dead):
        if dead is None:
            return None
        d = np.ascontiguousarray(dead, dtype=np.uint8)
        if d.size != self.graph.V + 1:
            raise ValueError("dead must have V+1 entries")
        return d

    def phase_trim1(self, dead=None) -> np.ndarray:
        d = self._dead(dead)
        out = np.zeros(self.graph.V + 1, np.uint8)
        _check(lib().scc_phase_trim1(self._ctx, _p8(d), _p8(out)))
        return out

    def phase_trim2(self, dead=None) -> np.ndarray:
        d = self._dead(dead)
        out = np.zeros(self.graph.V + 1, np.uint32)
        _check(lib().scc_phase_trim2(self._ctx, _p8(d), _p32(out)))
        return out

    def phase_first_pivot(self, dead=None) -> int:
        d = self._dead(dead)
        p = ctypes.c_uint32(0)
        _check(lib().scc_phase_first_pivot(self._ctx, _p8(d), ctypes.byref(p)))
        return p.value

    def phase_reach(self, src: int, direction: int = 0, expand: int = EXPAND_VERTEX, warp: int = 1,
                    dead=None) -> tuple[np.ndarray, int, float]:
        d = self._dead(dead)
        out = np.zeros(self.graph.V + 1, np.uint8)
        lv = ctypes.c_uint32(0)
        ms = ctypes.c_double(0)
        _check(lib().scc_phase_reach(self._ctx, ctypes.c_uint32(src), ctypes.c_int(direction),
                                     ctypes.c_int(expand), ctypes.c_int(warp), _p8(d), _p8(out),
                                     ctypes.byref(lv), ctypes.byref(ms)))
        return out, lv.value, ms.value

    def phase_min_label(self, dead=None) -> np.ndarray:
        d = self._dead(dead)
        out = np.zeros(self.graph.V + 1, np.uint32)
        _check(lib().scc_phase_min_label(self._ctx, _p8(d), _p32(out)))
        return out

    def phase_local(self) -> np.ndarray:
        """Local-SCC labels of detectSCC for the uploaded graph + partition."""
        out = np.zeros(self.graph.V + 1, np.uint32)
        _check(lib().scc_phase_local(self._ctx, _p32(out)))
        return out

    def phase_max_color(self, dead=None) -> np.ndarray:
        d = self._dead(dead)
        out = np.zeros(self.graph.V + 1, np.uint32)
        _check(lib().scc_phase_max_color(self._ctx, _p8(d), _p32(out)))
        return out


def scc(g: Graph, algo: str = "g", trim1: bool = True, trim2: bool = True, warp: int = 1,
        expand: int = EXPAND_AUTO, device: int = 0) -> tuple[np.ndarray, Stats]:
    """One-shot SCC decomposition: canonical labels (V+1) and run statistics."""
    with Context(device) as ctx:
        return ctx.solve(g, algo=algo, trim1=trim1, trim2=trim2, warp=warp, expand=expand)


# -- drop-in driver names (include/scc.h:49-53); they print the reference's lines --
def _dropin(name, *args):
    """Calls the C-linkage driver scc_<name>; returns its status (0 = ok).
    A failure's message is printed, as the reference's main prints what the
    drivers throw (src/scc.cu:104-108)."""
    rc = getattr(lib(), "scc_" + name)(*args)
    if rc != 0:
        print(lib().scc_last_error().decode(), flush=True)
    return rc


def vHong(CSize, RSize, Fc, Fr, Bc, Br, t1, t2) -> int:
    return _dropin("vHong", ctypes.c_uint32(CSize), ctypes.c_uint32(RSize), _p32(Fc), _p32(Fr),
                   _p32(Bc), _p32(Br), ctypes.c_bool(t1), ctypes.c_bool(t2))


def wHong(CSize, RSize, Fc, Fr, Bc, Br, t1, t2, warpSize) -> int:
    return _dropin("wHong", ctypes.c_uint32(CSize), ctypes.c_uint32(RSize), _p32(Fc), _p32(Fr),
                   _p32(Bc), _p32(Br), ctypes.c_bool(t1), ctypes.c_bool(t2), ctypes.c_int(warpSize))


def vSlota(CSize, RSize, Fc, Fr, Bc, Br, t1, t2) -> int:
    return _dropin("vSlota", ctypes.c_uint32(CSize), ctypes.c_uint32(RSize), _p32(Fc), _p32(Fr),
                   _p32(Bc), _p32(Br), ctypes.c_bool(t1), ctypes.c_bool(t2))


def wSlota(CSize, RSize, Fc, Fr, Bc, Br, t1, t2, warpSize) -> int:
    return _dropin("wSlota", ctypes.c_uint32(CSize), ctypes.c_uint32(RSize), _p32(Fc), _p32(Fr),
                   _p32(Bc), _p32(Br), ctypes.c_bool(t1), ctypes.c_bool(t2), ctypes.c_int(warpSize))


def detectSCC(CSize, RSize, Fc, Fr, Bc, Br, Pr, t1, t2) -> int:
    return _dropin("detectSCC", ctypes.c_uint32(CSize), ctypes.c_uint32(RSize), _p32(Fc), _p32(Fr),
                   _p32(Bc), _p32(Br), _p32(Pr), ctypes.c_bool(t1), ctypes.c_bool(t2))
