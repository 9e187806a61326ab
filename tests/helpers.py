"""Test helpers: digests, random graphs, label checks."""
import hashlib

import numpy as np

import detectingscc_amd as scc


def digest(lab: np.ndarray) -> str:
    """First 16 hex chars of SHA-256 over 'label(v)\\n' for v = 1..V."""
    text = "".join(f"{int(x)}\n" for x in lab[1:])
    return hashlib.sha256(text.encode()).hexdigest()[:16]


def random_graph(V: int, E: int, rng: np.random.Generator, self_loops: bool = True) -> scc.Graph:
    src = rng.integers(1, V + 1, size=E, dtype=np.uint32)
    dst = rng.integers(1, V + 1, size=E, dtype=np.uint32)
    if not self_loops:
        keep = src != dst
        src, dst = src[keep], dst[keep]
    return scc.from_arcs(V, src, dst)


def sizes_of(lab: np.ndarray) -> np.ndarray:
    s = np.bincount(lab[1:].astype(np.int64), minlength=lab.size)
    return s[s > 0]


def assert_same_labels(got: np.ndarray, want: np.ndarray, what: str = "") -> None:
    assert got.shape == want.shape, f"{what}: shape {got.shape} != {want.shape}"
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, (f"{what}: {bad.size} labels differ, first at v={bad[0]}: "
                           f"got {got[bad[0]]} want {want[bad[0]]}")
