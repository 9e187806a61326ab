// scc_dev.h -- device-side data layout of the SCC engine and the host
// launchers of its kernels.
//
// Per-vertex state in HBM (vertices 1..V, slot 0 is the dummy vertex):
//   cls[v]  uint32  class id of a live vertex; DEAD (0xFFFFFFFF) once v's SCC
//                   is known.  After a FW-BW split the ids are
//                   3*pivot + (fw + 2*bw), after "WCC" 3*min_reachable_id,
//                   during colouring the colour (a vertex id).  A neighbour
//                   probe "live and in my class" is one compare.
//   lab[v]  uint32  representative of v's SCC once finished (pivot id, own id
//                   for trim-1, pair minimum for trim-2); canonicalised to the
//                   minimum vertex id at the end.  A live vertex keeps the
//                   v | LAB_FINAL it starts with, so trim-1 writes no label.
//   aux[v]  uint32  scratch: colour during colouring, min id per label at the
//                   end.
//   fvis/bvis       one bit per vertex per direction; bits of finished
//                   vertices are pre-set at the start of a search, so a
//                   probe of a finished vertex never reads cls.
//   piv[c]  uint64  per-class pivot key max(min(outdeg*indeg, 2^32-1) << 32 |
//                   (2^32-1-v)), indexed by class id (3(V+1)+3 entries).
//   q*              frontier queues, V+1 entries each, two per direction.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sccamd {

constexpr uint32_t DEAD = 0xFFFFFFFFu;
// lab[v] flag: the label is already the canonical minimum of v's SCC (trim-1
// singletons, trim-2 pairs), so the final min-per-SCC pass skips it.
constexpr uint32_t LAB_FINAL = 0x80000000u;
constexpr int SWEEP_BLOCK = 256;
constexpr uint32_t LB_BIG = 512;   // degree from which a frontier vertex is a hub (LB mode)
constexpr uint32_t VIS_BLOCKS = 4096;   // most workgroups of the bitmap initialisation

// Visited bitmaps (fvis, bvis) keep two words per group of 32 vertices: the
// first has a bit for every visited or finished vertex, the second one for
// every finished vertex, so a bottom-up parent probe reads both with one
// 64-bit load.
__host__ __device__ __forceinline__ uint32_t vword(uint32_t v) { return (v >> 5) << 1; }

// Counters block: counter i lives at word CT(i), one 128-byte line per
// counter, so atomics on different counters never share a line (same-line
// atomics from every workgroup serialise on one L2 channel).  64-bit
// counters use words CT(i) and CT(i)+1.
constexpr int CS = 32;
__host__ __device__ constexpr int CT(int i) { return i * CS; }

enum : int {
    C_FWCNT = 0,     // 3 rotating FW queue counts
    C_BWCNT = 3,     // 3 rotating BW queue counts
    C_CHANGED = 6,   // fixpoint flag
    C_LIVE = 7,      // live vertices after an update
    C_TRIM1 = 8,     // vertices removed by trim-1
    C_TRIM2 = 9,     // pairs removed by trim-2
    C_PIVOT = 10,    // SCCs found by pivot / colour-root searches
    C_NPIV = 11,     // pivots elected this round
    C_NSCC = 12,     // canonical roots
    C_LARGEST = 13,  // largest SCC size
    C_FWV = 14,      // first-search statistics: |R| and |E_R| (64-bit pairs)
    C_FWE = 16,
    C_BWV = 18,
    C_BWE = 20,
    C_BIGF = 22,     // 3 rotating 64-bit packed (items, chunks) counts of the FW hub lists (22..24)
    C_BIGB = 25,     // same for BW (25..27)
    C_NLEVELS = 28,  // levels run by the narrow-frontier kernel, word 0 FW, word 1 BW
    C_PIV0 = 30,     // first pivot of the latest election
    C_ANOM = 31,     // invariant violations caught by kernel guards (must stay 0)
    C_NLIVE = 32,    // live vertices when the visited bitmaps were initialised
    C_TLIST = 33,    // 3 rotating counts of trim-1 removal lists (33..35)
    C_DOST = 36,     // direction-optimised search state: 2 slots x 2 directions (36..39)
    C_WORK = 40,     // work units of the first search, counted with sizes (40..47, 64-bit each)
    C_COUNT = 48
};

// Spread counters.  Same-address atomics serialise (about 12 ns each), so a
// counter that every workgroup of a sweep adds to, or reserves list space
// from, has kSpread slots besides its own word, each on a line of its own
// after the C_COUNT counter lines: workgroup b uses slot b % kSpread.  Its
// value is its own word plus its slots (the counter publication folds them
// for the host).  The spread counters are C_CHANGED..C_NSCC, the trim-1
// list counts C_TLIST..C_TLIST+2 and the search queue counts
// C_FWCNT..C_BWCNT+2 (the slots of each direction's three counts are
// consecutive: count k's slot j is CSLOT(first, k * kSpread + j)).
constexpr int kSpread = 32;
__host__ __device__ constexpr int spread_index(int c)
{
    return (c >= C_CHANGED && c <= C_NSCC) ? c - C_CHANGED
           : (c >= C_TLIST && c < C_TLIST + 3) ? C_NSCC - C_CHANGED + 1 + (c - C_TLIST)
           : (c >= C_FWCNT && c < C_BWCNT + 3) ? C_NSCC - C_CHANGED + 4 + (c - C_FWCNT)
                                                : -1;
}
constexpr int kSpreadCounters = C_NSCC - C_CHANGED + 1 + 3 + 6;
constexpr int kCtrLines = C_COUNT + kSpreadCounters * kSpread;   // lines of the counters block
// word of slot j of spread counter c
__host__ __device__ constexpr int CSLOT(int c, int j) { return CT(C_COUNT + spread_index(c) * kSpread + j); }

// A list whose count is a spread counter has kSpread + 1 segments: segment
// 0 from index 0 (entries reserved from the counter's own word), segment
// j + 1 from index j * cap (entries reserved from slot j, at most cap).
// One pass writes a list one way only (a trim-1 sweep or a bottom-up level
// through the slots, a trim-1 list pass, a top-down level or the seeding
// through the own word), so segment 0 and slot 0's region never both hold
// entries.
// seg_prefix writes pre[k] = the entries of the segments before k
// (k = 0..kSpread+1, so pre[kSpread + 1] is the list's length), given the
// counter's own word and its slot 0 (slot j is CT(j) words further on).
// Called by the whole workgroup; pre is in LDS.
__device__ __forceinline__ void seg_prefix(const uint32_t *own, const uint32_t *slot0, uint32_t cap, uint32_t *pre)
{
    if (threadIdx.x < 64) {
        const uint32_t l = threadIdx.x;
        uint32_t x = 0;
        if (l == 0) x = *own;
        else if (l <= (uint32_t)kSpread) x = min(slot0[CT((int)l - 1)], cap);
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if ((int)l >= o) x += y;
        }
        if (l <= (uint32_t)kSpread) pre[l + 1] = x;
        if (l == 0) pre[0] = 0;
    }
    __syncthreads();
}

// Index in the list of logical entry i < pre[kSpread + 1]: the last
// segment k whose prefix is <= i holds it.
__device__ __forceinline__ uint32_t seg_pos(const uint32_t *pre, uint32_t i, uint32_t cap)
{
    uint32_t lo = 0, hi = kSpread;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo == 0 ? i : (lo - 1) * cap + (i - pre[lo]);
}

// Clears a spread counter (its own word and slots); threads 0..kSpread of
// one workgroup call it.
__device__ __forceinline__ void seg_clear(uint32_t *own, uint32_t *slot0)
{
    if (threadIdx.x == 0) *own = 0;
    else if (threadIdx.x <= (uint32_t)kSpread) slot0[CT((int)threadIdx.x - 1)] = 0;
}

struct DevGraph {
    uint32_t V = 0, E = 0;
    uint32_t *Fr = nullptr, *Fc = nullptr, *Br = nullptr, *Bc = nullptr;
    // F1[v] = Fc[Fr[v]] and B1[v] = Bc[Br[v]] (0 for an empty row): the first
    // out- and in-neighbour of every vertex in one coalesced word, for the
    // probes that read only a list's first entry (trim-1, bottom-up), which
    // would otherwise fetch a 64-byte line at a random address per vertex
    uint32_t *F1 = nullptr, *B1 = nullptr;
    // DP[v] = min(outdeg(v) * indeg(v), 2^32 - 1): the high word of v's pivot
    // key, one word per vertex instead of four row offsets
    uint32_t *DP = nullptr;
};

struct DevState {
    uint32_t *cls = nullptr, *lab = nullptr, *aux = nullptr;
    uint32_t *fvis = nullptr, *bvis = nullptr;
    uint64_t *piv = nullptr;
    uint32_t *qf[2] = {nullptr, nullptr}, *qb[2] = {nullptr, nullptr};
    uint32_t *pivlist = nullptr;   // pivots of the current round
    uint32_t *stamp = nullptr;     // trim-1 list passes: epoch of the last test of each vertex
    uint32_t *bigf[3] = {}, *bigb[3] = {};           // hub lists, rotating like the counts
    uint32_t *bigoff_f[3] = {}, *bigoff_b[3] = {};
    uint32_t *part = nullptr;      // partition ids (detectSCC)
    uint32_t *ctr = nullptr;       // kCtrLines lines (CT(kCtrLines) words)
    uint64_t *bred = nullptr;      // per-workgroup (live count, best key) of the bitmap initialisation
    uint32_t nwords = 0;           // words of a one-bit-per-vertex bitmap (fvis/bvis hold twice that)
};

// One search direction of a BFS level.
struct DirArgs {
    const uint32_t *R, *C;       // CSR of this direction
    uint32_t *vis;               // visited bitmap
    const uint32_t *qin;
    uint32_t *qout;
    uint32_t *cnt;               // the 3 rotating counts (spread counters)
    uint32_t *cslot;             // their slots: count k's slot j at cslot[CT(k * kSpread + j)]
    // LB mode defers each hub (degree >= LB_BIG) of level L to a list whose
    // chunks the level L+1 kernel expands before its own tiles, so a level
    // is one launch; three lists rotate like the counts
    uint32_t *big;               // list this level appends to
    uint32_t *bigoff;            // chunk offset of each hub
    unsigned long long *bigc;    // its packed (items << 32 | chunks)
    const uint32_t *pbig;        // previous level's list, expanded by this level
    const uint32_t *pbigoff;
    const unsigned long long *pbigc;
    unsigned long long *bigc_zero;   // the list the next level appends to, cleared here
};

// A direction-optimised level (top-down or bottom-up per direction, chosen
// on the device).
struct DoArgs {
    DirArgs d[2];
    const uint32_t *PR[2], *PC[2];   // parent lists of each direction (the opposite CSR)
    const uint32_t *PF[2];           // first parent of each vertex (the opposite CSR's F1/B1)
    const uint32_t *cls;
    uint32_t *ctr;                   // counters: search state (C_DOST), C_NLIVE, C_NPIV
    uint32_t V;
    int ndirs;
    int level;
    int bu_ok;                       // 0 on a grid too small for a bottom-up pass: stay top-down
    uint32_t qcap;                   // slot region of a bottom-up level's queue: slot_cap(V, full grid)
    int count;                       // 1: add the level's work units (WorkCount) to C_WORK..C_WORK+7
};

// Work units of a direction-optimised search, counted per lane in the
// untimed run with sizes and added to the counters C_WORK + k (64-bit) at
// the end of each level.  Their bytes (the search's algorithmic traffic):
//   td_vertices  12 B  queue entry + row offsets of an expanded vertex
//   td_arcs       8 B  column entry + visited word of a scanned arc
//   claims        4 B  queue entry of a newly visited vertex
//   bu_swept    1/8 B  visited bit of a vertex a bottom-up pass looks at
//   bu_cand       4 B  first-parent entry of an unvisited candidate
//   bu_first      8 B  visited/finished word of that first parent
//   bu_rows       8 B  row offsets of a candidate whose first parent missed
//   bu_more      12 B  parent id + visited/finished word of a further parent
struct WorkCount {
    uint32_t n[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};
enum : int { W_TDV = 0, W_TDE, W_CLAIM, W_BUS, W_BUC, W_BUF, W_BUR, W_BUM };

// Direction-optimising switch points (Beamer et al.): go bottom-up when the
// frontier holds more than 1/kToBottomUp of the vertices still unvisited,
// or more than 1/kToBottomUpBurst of them right after growing kBurst-fold
// in one level (a frontier that bursts -- the neighbours of a hub -- has a
// high average degree, which the vertex counts alone do not show), and back
// to top-down when it falls under 1/kToTopDown of the live vertices.  A
// bottom-up pass reads every vertex's bitmap word, so a frontier under
// 1/kBottomUpFloor of the live vertices stays top-down even when few
// vertices remain unvisited (the tail of a search).
constexpr int64_t kToBottomUp = 16;
constexpr int64_t kToBottomUpBurst = 128;
constexpr int64_t kBurst = 4;
constexpr int64_t kToTopDown = 48;
constexpr int64_t kBottomUpFloor = 256;

enum : uint32_t { DIR_OFF = 0, DIR_TD = 1, DIR_BU = 2 };

// Search state of one direction: the mode its level runs in, the level at
// which it finished, the frontier it started from and the vertices it has
// visited.  Held in counter line C_DOST + 2 * slot + dir (words 0..5).
struct DirState {
    uint32_t mode, end, nf, pad;
    uint64_t visited;
};
constexpr int kDirStateWords = 6;

// The state of level L from that of level L-1, given the vertices level L-1
// queued (nf), the hubs it deferred to level L (top-down only: a bottom-up
// pass defers nothing), the live vertices of the search and whether level
// L may run bottom-up.
__host__ __device__ __forceinline__ DirState dir_decide(DirState s, uint32_t nf, uint32_t hubs, uint32_t nlive,
                                                        uint32_t level, bool bu_ok)
{
    if (s.mode == DIR_OFF) return s;
    if (s.mode == DIR_BU) hubs = 0;
    if (!nf && !hubs) {
        s.mode = DIR_OFF;
        s.end = level;
        return s;
    }
    const int64_t grew = (int64_t)nf > kBurst * (int64_t)s.nf;
    s.nf = nf;
    s.visited += nf;
    const int64_t remaining = (int64_t)nlive - (int64_t)s.visited;
    if (s.mode == DIR_TD && bu_ok && (int64_t)nf * (grew ? kToBottomUpBurst : kToBottomUp) > remaining &&
        (int64_t)nf * kBottomUpFloor >= (int64_t)nlive)
        s.mode = DIR_BU;
    else if (s.mode == DIR_BU && (int64_t)nf * kToTopDown < (int64_t)nlive) s.mode = DIR_TD;
    return s;
}

__host__ __device__ __forceinline__ DirState load_dir_state(const uint32_t *ctr, int slot, int dir)
{
    const uint32_t *w = ctr + CT(C_DOST + 2 * slot + dir);
    DirState s;
    s.mode = w[0];
    s.end = w[1];
    s.nf = w[2];
    s.pad = 0;
    s.visited = (uint64_t)w[4] | ((uint64_t)w[5] << 32);
    return s;
}

__device__ __forceinline__ void store_dir_state(uint32_t *ctr, int slot, int dir, const DirState &s)
{
    uint32_t *w = ctr + CT(C_DOST + 2 * slot + dir);
    w[0] = s.mode;
    w[1] = s.end;
    w[2] = s.nf;
    w[3] = 0;
    w[4] = (uint32_t)s.visited;
    w[5] = (uint32_t)(s.visited >> 32);
}

struct LevelArgs {
    DirArgs d[2];
    const uint32_t *cls;
    int ndirs;                   // 1 = FW only, 2 = FW and BW
    int level;                   // selects the rotating counters
};

// Two-direction launches (grid = 2 * bpd).  Workgroups are dispatched to the
// eight XCDs round-robin (workgroup x runs on XCD x mod 8), so taking the
// direction from bit 2 of x runs direction 0 on XCDs 0-3 and direction 1 on
// XCDs 4-7: each XCD's L2 then caches one direction's bitmap instead of
// both.  Needs bpd % 4 == 0 when ndirs == 2 (dir_blocks rounds it up).
__device__ __forceinline__ void split_dirs(int ndirs, uint32_t bpd, int &dir, uint32_t &bid)
{
    const uint32_t x = blockIdx.x;
    if (ndirs == 2) {
        dir = (int)((x >> 2) & 1u);
        bid = ((x >> 3) << 2) | (x & 3u);
    } else {
        dir = 0;
        bid = x;
    }
    (void)bpd;
}
__host__ __forceinline__ uint32_t dir_blocks(uint32_t bpd, int ndirs)
{
    return ndirs == 2 ? (bpd + 3u) & ~3u : bpd;
}

// ---- sweep kernels (kernels_sweep.hip) ----
void k_fill_u32(hipStream_t s, uint32_t *p, uint32_t val, uint64_t n);
// g.F1, g.B1 and g.DP from the CSR (after an upload)
void k_first_arcs(hipStream_t s, const DevGraph &g);
void k_init_state(hipStream_t s, const DevGraph &g, DevState &st, const uint8_t *dead_dev,
                  bool vertices = true);
// trim-1 pass `pass` of a fixpoint: pass 0 (and a pass whose input list,
// counted at C_TLIST + pass%3, holds more than V / sweep_share vertices)
// sweeps every vertex, the others test the live neighbours of the vertices
// listed in `in`; removals are listed in `out`, counted at
// C_TLIST + (pass+1)%3.  A sweep lists workgroup b's removals in the region
// of slot b % kSpread (slot j's region starts at j * trim_slot_cap(V)) and
// counts them in that slot; a list pass lists from index 0 and counts in
// the counter's own word.  `in` and `out` hold trim_list_words(V) words.
// `epoch` is new for every pass of the context
// (st.stamp never holds it).  With `live` (only while every live vertex has
// class 0), neighbour tests read a live bitmap there instead of classes;
// pass 0 builds it from the class words.  With degree_first (a fixpoint
// from a fresh state: no vertex dead, all in class 0, labels v | LAB_FINAL)
// pass 0 removes only vertices without an in- or out-arc, since every
// neighbour is live then (it reads each vertex's degree product), and
// writes every vertex's class word and label itself (unless `labels` is
// false), so the fresh state's per-vertex words need not be written before.
void k_trim1(hipStream_t s, const DevGraph &g, DevState &st, const uint32_t *in, uint32_t *out,
             uint32_t epoch, uint32_t pass, uint32_t sweep_share, uint32_t *live, bool degree_first,
             bool labels = true);   // labels = false (single class only): no label or class word is
                                    // written, the live bitmap alone tracks removals
// Largest number of vertices the workgroups b with b % kSpread == j own in
// a pass over 64-vertex chunks where wave w of `grid` 4-wave workgroups
// takes chunks w, w + 4 * grid, ...: the size of slot j's list region.
uint32_t slot_cap(uint32_t V, uint32_t grid);
uint32_t trim_slot_cap(uint32_t V);
size_t trim_list_words(uint32_t V);
// live_bits: every live vertex is in class 0 and the bitmap (as the
// single-class trim-1 leaves it) is exact; tested instead of the class words
void k_trim2(hipStream_t s, const DevGraph &g, DevState &st, const uint32_t *live_bits = nullptr);
// Visited bitmaps of a search and C_NLIVE; with elect_first (every live
// vertex in class 0) also the first pivot, seeded in both directions.
// live_bits: every live vertex is in class 0 and the bitmap (1 bit per
// vertex, as the single-class trim-1 leaves it) is exact; read instead of
// the class words.
void k_init_vis(hipStream_t s, const DevGraph &g, DevState &st, bool elect_first,
                const uint32_t *live_bits = nullptr);
// single_class: every live vertex is in class 0 (the first pivot)
void k_elect(hipStream_t s, const DevGraph &g, DevState &st);
void k_select(hipStream_t s, const DevGraph &g, DevState &st, int seed_bw);
// scc_min: a round with one pivot (C_PIV0) also leaves its SCC's smallest
// vertex in aux[pivot]
// live_bits as for k_init_vis; lazy (a first round after a label-less
// degree pass): the members of the pivot's SCC keep their words, which
// k_label_from_bits (nothing left live) or k_label_fixup writes
void k_update(hipStream_t s, const DevGraph &g, DevState &st, bool scc_min = false,
              const uint32_t *live_bits = nullptr, bool lazy = false);
// The lazy first round's update (one class, one pivot C_PIV0, live_bits
// exact) a 32-vertex group per lane; same results as k_update(scc_min,
// live_bits, lazy) including the pivot-entry clearing.
void k_update_words(hipStream_t s, const DevGraph &g, DevState &st, const uint32_t *live_bits);
// Canonical labels of a decomposition that a lazy first round finished, from
// the search's bitmaps (gated: a no-op unless C_LIVE is 0).
void k_label_from_bits(hipStream_t s, const DevGraph &g, DevState &st);
// After a lazy first round that left live vertices: the labels and class
// words an eager first round would have left (live_bits: the trim's bitmap).
void k_label_fixup(hipStream_t s, const DevGraph &g, DevState &st, const uint32_t *live_bits);
void k_clear_piv(hipStream_t s, DevState &st, uint32_t max_pivots);
void k_reach_stats(hipStream_t s, const DevGraph &g, DevState &st);
void k_minlabel_init(hipStream_t s, const DevGraph &g, DevState &st);
void k_minlabel_prop(hipStream_t s, const DevGraph &g, DevState &st);
void k_minlabel_finish(hipStream_t s, const DevGraph &g, DevState &st);
void k_color_init(hipStream_t s, const DevGraph &g, DevState &st);
void k_color_prop(hipStream_t s, const DevGraph &g, DevState &st);
void k_color_roots(hipStream_t s, const DevGraph &g, DevState &st);
void k_color_finish(hipStream_t s, const DevGraph &g, DevState &st);
// Canonical labels; gated: each pass is a no-op unless C_LIVE is 0 (queued
// before the host has read the live count of the round before).
void k_canon(hipStream_t s, const DevGraph &g, DevState &st, bool sizes, bool gated = false);
// SCC sizes from canonical labels into the class-word array (root r's SCC
// has cnt[r] + 1 vertices) and the largest size into C_LARGEST
void k_sizes(hipStream_t s, const DevGraph &g, DevState &st, bool gated = false);
// Size histogram after k_sizes: (size, count) pairs, unsorted, into pairs
// (up to cap), their number into *npairs (zeroed by the caller); hist is
// scratch of V + 1 words
void k_size_histogram(hipStream_t s, const DevGraph &g, DevState &st, uint32_t *hist, uint64_t *pairs,
                      uint32_t cap, uint32_t *npairs);
// The apply pass alone: every non-final label's minimum is already in aux
// (k_update with scc_min in a round that left no live vertex).
void k_canon_apply(hipStream_t s, const DevGraph &g, DevState &st, bool gated);
void k_seed_single(hipStream_t s, DevState &st, uint32_t src, int dir);
void k_export_visited(hipStream_t s, const DevGraph &g, DevState &st, int dir, uint8_t *out_dev);
void k_export_removed(hipStream_t s, const DevGraph &g, DevState &st, uint8_t *out_dev);

// ---- reachability (kernels_reach.hip) ----
// expand: SCC_EXPAND_VERTEX / VWARP / LB; warp = lanes per vertex for VWARP
void k_expand_level(hipStream_t s, const LevelArgs &a, int expand, int warp, bool single_class,
                    uint32_t blocks_per_dir);
// One bottom-up pass of each listed direction: every live unvisited vertex
// scans its parents (PR/PC = the opposite CSR) for a visited one of its class.
// Direction-optimised search: k_do_begin sets slot 0 (the first `ndirs`
// directions top-down, visited = the seeds), then one k_do_level per level.
void k_do_begin(hipStream_t s, uint32_t *ctr, int ndirs);
void k_do_level(hipStream_t s, const DoArgs &a, bool single_class, uint32_t blocks_per_dir);
// Narrow-frontier levels (one workgroup per direction) from level a.level,
// leaving that level's input rewritten; a no-op unless that level would run
// top-down from a small frontier with no hubs pending.  narrow_cap(): the
// largest frontier the host starts it on
void k_narrow_levels(hipStream_t s, const DoArgs &a, bool single_class);
uint32_t narrow_cap();
uint32_t do_blocks(uint32_t V, int ndirs);
// Words of a search queue: bottom-up levels write kSpread slot regions.
size_t queue_words(uint32_t V);

uint32_t sweep_blocks(uint32_t V);

} // namespace sccamd
