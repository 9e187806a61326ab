"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, as the checker or the CPU baseline.  The
product never imports it.  Graph arguments are any object with V, E and
uint32 numpy arrays Fr, Fc, Br, Bc in the reference layout.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from types import SimpleNamespace

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_L = None


class _Stats(ctypes.Structure):
    _fields_ = [("num_sccs", ctypes.c_int64), ("num_trim1", ctypes.c_int64),
                ("num_trim2", ctypes.c_int64), ("num_pivot", ctypes.c_int64),
                ("rounds", ctypes.c_int64)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for f in ["orc_tarjan", "orc_kosaraju", "orc_brute", "orc_trim1_fixpoint",
                  "orc_trim2_snapshot", "orc_reach", "orc_local_scc", "orc_vhong", "orc_vslota"]:
            getattr(L, f).restype = ctypes.c_int64
        L.orc_first_pivot.restype = ctypes.c_uint32
        _L = L
    return _L


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(_u8p if a.dtype == np.uint8 else _u32p)


def _u8(a, n):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.uint8)
    assert a.size == n
    return a


def load_gtgraph(path: str):
    L = lib()
    V, E = ctypes.c_uint32(), ctypes.c_uint32()
    ptrs = [_u32p() for _ in range(4)]
    r = L.orc_load_gtgraph(os.fsencode(path), ctypes.byref(V), ctypes.byref(E),
                           *[ctypes.byref(p) for p in ptrs])
    if r != 0:
        raise IOError(f"oracle loader failed ({r}) on {path}")
    V, E = V.value, E.value
    sizes = [V + 2, E, V + 2, E]
    arrs = []
    for p, n in zip(ptrs, sizes):
        arrs.append(np.ctypeslib.as_array(p, shape=(max(n, 1),))[:n].copy())
        L.orc_free(p)
    return SimpleNamespace(V=V, E=E, Fr=arrs[0], Fc=arrs[1], Br=arrs[2], Bc=arrs[3])


def build_csr(V: int, src, dst):
    s = np.ascontiguousarray(src, dtype=np.uint32)
    d = np.ascontiguousarray(dst, dtype=np.uint32)
    E = s.size
    Fr = np.zeros(V + 2, np.uint32); Br = np.zeros(V + 2, np.uint32)
    Fc = np.zeros(max(E, 1), np.uint32); Bc = np.zeros(max(E, 1), np.uint32)
    lib().orc_build_csr(ctypes.c_uint32(V), ctypes.c_uint32(E), _p(s), _p(d), _p(Fr), _p(Fc), _p(Br), _p(Bc))
    return SimpleNamespace(V=V, E=E, Fr=Fr, Fc=Fc[:E], Br=Br, Bc=Bc[:E])


def load_partition(path: str, V: int) -> np.ndarray:
    part = np.zeros(V + 2, np.uint32)
    if lib().orc_load_partition(os.fsencode(path), ctypes.c_uint32(V), _p(part)) != 0:
        raise IOError(f"oracle partition reader failed on {path}")
    return part


def _g(g):
    return ctypes.c_uint32(g.V), _p(g.Fr), _p(g.Fc), _p(g.Br), _p(g.Bc)


def tarjan(g) -> tuple[np.ndarray, int]:
    lab = np.zeros(g.V + 1, np.uint32)
    n = lib().orc_tarjan(ctypes.c_uint32(g.V), _p(g.Fr), _p(g.Fc), _p(lab))
    return lab, int(n)


def kosaraju(g) -> tuple[np.ndarray, int]:
    lab = np.zeros(g.V + 1, np.uint32)
    n = lib().orc_kosaraju(*_g(g), _p(lab))
    return lab, int(n)


def brute(g) -> tuple[np.ndarray, int]:
    lab = np.zeros(g.V + 1, np.uint32)
    n = lib().orc_brute(ctypes.c_uint32(g.V), _p(g.Fr), _p(g.Fc), _p(lab))
    if n < 0:
        raise ValueError("brute force supports V <= 63")
    return lab, int(n)


def trim1(g, dead=None) -> np.ndarray:
    """Vertices removed by the trim-1 fixpoint, starting with `dead` removed."""
    d = _u8(dead, g.V + 1)
    out = np.zeros(g.V + 1, np.uint8)
    lib().orc_trim1_fixpoint(*_g(g), _p(d), _p(out))
    return out


def trim2_snapshot(g, dead) -> np.ndarray:
    d = _u8(dead, g.V + 1)
    out = np.zeros(g.V + 1, np.uint32)
    lib().orc_trim2_snapshot(*_g(g), _p(d), _p(out))
    return out


def first_pivot(g, dead=None) -> int:
    d = _u8(dead, g.V + 1)
    return int(lib().orc_first_pivot(ctypes.c_uint32(g.V), _p(g.Fr), _p(g.Br), _p(d)))


def reach(g, src: int, direction: int = 0, dead=None) -> np.ndarray:
    d = _u8(dead, g.V + 1)
    out = np.zeros(g.V + 1, np.uint8)
    R, C = (g.Fr, g.Fc) if direction == 0 else (g.Br, g.Bc)
    lib().orc_reach(ctypes.c_uint32(g.V), _p(R), _p(C), ctypes.c_uint32(src), _p(d), _p(out))
    return out


def min_label(g, dead=None) -> np.ndarray:
    d = _u8(dead, g.V + 1)
    out = np.zeros(g.V + 1, np.uint32)
    lib().orc_min_reach_label(ctypes.c_uint32(g.V), _p(g.Fr), _p(g.Fc), _p(d), _p(out))
    return out


def max_color(g, dead=None) -> np.ndarray:
    d = _u8(dead, g.V + 1)
    out = np.zeros(g.V + 1, np.uint32)
    lib().orc_max_reach_color(ctypes.c_uint32(g.V), _p(g.Fr), _p(g.Fc), _p(d), _p(out))
    return out


def local_scc(g, part) -> tuple[np.ndarray, int]:
    p = np.ascontiguousarray(part, dtype=np.uint32)
    lab = np.zeros(g.V + 1, np.uint32)
    n = lib().orc_local_scc(ctypes.c_uint32(g.V), _p(g.Fr), _p(g.Fc), _p(p), _p(lab))
    return lab, int(n)


def vhong(g, t1=True, t2=True, threads=0):
    lab = np.zeros(g.V + 1, np.uint32)
    st = _Stats()
    n = lib().orc_vhong(*_g(g), ctypes.c_int(int(t1)), ctypes.c_int(int(t2)), ctypes.c_int(threads),
                        _p(lab), ctypes.byref(st))
    return lab, int(n), st


def vslota(g, t1=True, threads=0):
    lab = np.zeros(g.V + 1, np.uint32)
    st = _Stats()
    n = lib().orc_vslota(*_g(g), ctypes.c_int(int(t1)), ctypes.c_int(threads), _p(lab), ctypes.byref(st))
    return lab, int(n), st


def max_threads() -> int:
    return int(lib().orc_max_threads())
