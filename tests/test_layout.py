"""Host-side emulation of the device lists' slot layout, with bounds checks.

A trim-1 sweep and a bottom-up search level write what they remove or visit
into the region of their workgroup's slot (workgroup b uses slot b % 32; a
slot's region starts at slot * cap), and the next pass reads the regions
through their prefix sums.  That is in bounds only if no slot's workgroups
own more vertices than `cap`, and the 32 regions fit the allocated list.
Here the kernels' chunk ownership (wave w = 4 b + lane-group owns the
64-vertex chunks c with c = w mod 4 * grid) is enumerated independently of
the library's capacity formulas, whose values come from the library
(scc_layout_info), at the benchmark sizes and at every size boundary."""
import ctypes

import numpy as np
import pytest

import detectingscc_amd as scc

SIZES = [1, 2, 63, 64, 65, 127, 1000, 65535, 65536, 65537, (1 << 20) - 1, 1 << 20, (1 << 20) + 1,
         (1 << 20) + 255, 4 << 20, 1 << 22, 1 << 23, 4096 * 4096, 1 << 24, (1 << 24) + 12345, 100_000_003,
         (1 << 30) - 1]


def layout(V):
    out = (ctypes.c_uint64 * 10)()
    assert scc.lib().scc_layout_info(ctypes.c_uint32(V), out, 10) == 0
    keys = ["sweep_blocks", "trim_slot_cap", "trim_list_words", "do_blocks1", "do_blocks2", "bu_cap1",
            "bu_cap2", "queue_words", "spread", "narrow_cap"]
    return dict(zip(keys, (int(x) for x in out)))


def per_slot(V, grid, spread):
    """Vertices (0..V) owned by each slot when `grid` workgroups of 4 waves
    sweep 64-vertex chunks, wave w owning chunks c with c % (4 grid) == w."""
    nchunks = V // 64 + 1
    ch = np.arange(nchunks, dtype=np.int64)
    size = np.minimum(64, V + 1 - ch * 64)           # vertices 0..V in chunk c
    block = (ch % (4 * grid)) // 4
    return np.bincount(block % spread, weights=size, minlength=spread).astype(np.int64)


@pytest.mark.parametrize("V", SIZES)
def test_trim_sweep_regions_in_bounds(V):
    L = layout(V)
    owned = per_slot(V, L["sweep_blocks"], L["spread"])
    assert owned.sum() == V + 1
    assert owned.max() <= L["trim_slot_cap"], (V, owned.max(), L["trim_slot_cap"])
    # the 32 regions, and a list pass (which writes from 0, each vertex once)
    assert L["spread"] * L["trim_slot_cap"] <= L["trim_list_words"]
    assert L["trim_list_words"] >= V + 1


@pytest.mark.parametrize("V", SIZES)
@pytest.mark.parametrize("ndirs", [1, 2])
def test_bottom_up_regions_in_bounds(V, ndirs):
    L = layout(V)
    bpd = L[f"do_blocks{ndirs}"]
    cap = L[f"bu_cap{ndirs}"]
    if ndirs == 2:
        assert bpd % 4 == 0       # the direction comes from bit 2 of the workgroup index
    owned = per_slot(V, bpd, L["spread"])
    assert owned.sum() == V + 1
    assert owned.max() <= cap, (V, ndirs, owned.max(), cap)
    assert L["spread"] * cap <= L["queue_words"]
    assert L["queue_words"] >= V + 1                  # a top-down level or the narrow kernel: <= V entries


@pytest.mark.parametrize("V", [1 << 24, 4096 * 4096])
def test_segment_positions(V):
    """seg_pos: list index i of a slotted list maps into slot j's region at
    j * cap + (i - pre[j]); with any per-slot counts up to what each slot
    owns, every position stays below the list's words (randomised counts,
    the pass counts of a trim-1 fixpoint on R-MAT scale 24 removing from a
    few vertices to millions)."""
    L = layout(V)
    rng = np.random.default_rng(V)
    owned = per_slot(V, L["sweep_blocks"], L["spread"])
    cap = L["trim_slot_cap"]
    for frac in [0.0, 1e-6, 0.001, 0.03, 0.3, 1.0]:
        counts = np.minimum(owned, rng.binomial(owned, frac))
        pre = np.concatenate([[0], np.cumsum(counts)])
        n = int(pre[-1])
        idx = np.unique(np.concatenate([[0, max(n - 1, 0)], rng.integers(0, max(n, 1), 1000)]))
        idx = idx[idx < n]
        seg = np.searchsorted(pre, idx, side="right") - 1
        pos = seg * cap + (idx - pre[seg])
        assert (pos < (seg + 1) * cap).all() and (pos < L["trim_list_words"]).all()
