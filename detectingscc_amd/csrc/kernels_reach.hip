// kernels_reach.hip -- one BFS level of forward / backward reachability
// (fwd/bwd scc_kernels.cu:360-445, fwd_warp/bwd_warp scc_kernels.h:46-105)
// rebuilt as data-driven frontier expansion for gfx950.
//
// A level reads the input queue of each direction and writes the next one.
// Claims are atomicOr on a visited bitmap (finished vertices pre-set), so a
// vertex enters a queue at most once per search.  Queue counts rotate
// through three device counters so that no host step sits between levels:
// level L reads cnt[L%3], appends to cnt[(L+1)%3] and zeroes cnt[(L+2)%3]
// for level L+1.  With `single_class` every live vertex is in one class and
// the class compare is skipped (the probe is one bitmap bit).
//
// Atomics on one address serialise (~10 ns each on MI355X), so the fast
// kernels (LB, bottom-up) stage pushes in LDS and reserve queue space
// with one global atomic per ~1,800 pushes; the reference-mapping kernels
// (vertex, virtual warp) aggregate per wave.
#include <algorithm>

#include "lds_queue.h"
#include "scc_dev.h"
#include "scc_amd.h"

namespace sccamd {

__device__ __forceinline__ uint32_t wlane() { return threadIdx.x & 63u; }

template <bool SINGLE>
__device__ __forceinline__ bool try_claim(uint32_t *vis, const uint32_t *cls, uint32_t w, uint32_t c)
{
    const uint32_t bit = 1u << (w & 31);
    if (vis[vword(w)] & bit) return false;
    if (!SINGLE && cls[w] != c) return false;
    return !(atomicOr(&vis[vword(w)], bit) & bit);
}

// Append w for every active lane with `claimed`; one atomic per wave.
__device__ __forceinline__ void push_wave(bool claimed, uint32_t w, uint32_t *qout, uint32_t *cnt)
{
    const uint64_t act = __ballot(1);
    const uint64_t m = __ballot(claimed);
    if (!m) return;
    const int leader = __ffsll((unsigned long long)act) - 1;
    uint32_t base = 0;
    if ((int)wlane() == leader) base = atomicAdd(cnt, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (claimed) qout[base + (uint32_t)__popcll(m & ((1ull << wlane()) - 1))] = w;
}

// W lanes per frontier vertex; W = 1 is the vertex-per-thread mapping of
// fwd/bwd, W > 1 the virtual warp of fwd_warp<W> (lanes stride the list).
// A frontier vertex of degree >= LB_BIG (a hub: one lane, or one virtual
// warp, would walk its whole list while the rest of the grid idles) is
// instead listed for the next level, whose launch first cuts the listed
// hubs into LB_CHUNK-arc chunks over all its lanes (the load-balanced
// mapping's hub lists).  The hub's neighbours join one level late, which
// leaves every reachable set unchanged.
constexpr uint32_t VW_HUB = LB_BIG;   // the hub lists hold at most E / LB_BIG entries
constexpr uint32_t VW_CHUNK = 512;   // LB_CHUNK

template <int W, bool SINGLE>
__global__ __launch_bounds__(256) void expand_vw_kernel(LevelArgs a, uint32_t bpd)
{
    int dir;
    uint32_t bid;
    split_dirs(a.ndirs, bpd, dir, bid);
    const DirArgs &D = a.d[dir];
    const int L = a.level;
    // only bottom-up levels (direction-optimised searches) write the counts'
    // slots, and a search starts with them clear: this mapping's counts are
    // the counters' own words
    const uint32_t n = D.cnt[CT(L % 3)];
    uint32_t *cnt_out = &D.cnt[CT((L + 1) % 3)];
    if (bid == 0) {
        seg_clear(&D.cnt[CT((L + 2) % 3)], &D.cslot[CT(((L + 2) % 3) * kSpread)]);
        if (threadIdx.x == 0) *D.bigc_zero = 0ull;
    }
    const uint32_t *__restrict__ R = D.R;
    const uint32_t *__restrict__ C = D.C;
    // hubs of the previous level: chunk gch belongs to the item whose chunk
    // offset is the last one <= gch
    const unsigned long long packed = *D.pbigc;
    const uint32_t items = (uint32_t)(packed >> 32), chunks = (uint32_t)packed;
    for (uint32_t gch = bid; gch < chunks; gch += bpd) {
        uint32_t lo = 0, hi = items - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (D.pbigoff[mid] <= gch) lo = mid;
            else hi = mid - 1;
        }
        const uint32_t v = D.pbig[lo];
        const uint32_t c = SINGLE ? 0u : a.cls[v];
        const uint32_t b = R[v] + (gch - D.pbigoff[lo]) * VW_CHUNK;
        const uint32_t e = min(R[v + 1], b + VW_CHUNK);
        for (uint32_t j = b + threadIdx.x; j < e; j += blockDim.x) {
            const uint32_t w = C[j];
            push_wave(try_claim<SINGLE>(D.vis, a.cls, w, c), w, D.qout, cnt_out);
        }
    }
    if (n == 0) return;
    const uint32_t g = (bid * blockDim.x + threadIdx.x) / W;
    const uint32_t ng = bpd * blockDim.x / W;
    const uint32_t sub = threadIdx.x % W;
    for (uint32_t i = g; i < n; i += ng) {
        const uint32_t v = D.qin[i];
        const uint32_t c = SINGLE ? 0u : a.cls[v];
        const uint32_t b = R[v], e = R[v + 1];
        if (e - b >= VW_HUB) {
            if (sub == 0) {
                const unsigned long long nch = (e - b + VW_CHUNK - 1) / VW_CHUNK;
                const unsigned long long old = atomicAdd(D.bigc, (1ull << 32) | nch);
                const uint32_t item = (uint32_t)(old >> 32);
                D.big[item] = v;
                D.bigoff[item] = (uint32_t)old;
            }
            continue;
        }
        for (uint32_t j = b + sub; j < e; j += W) {
            const uint32_t w = C[j];
            const bool claimed = try_claim<SINGLE>(D.vis, a.cls, w, c);
            push_wave(claimed, w, D.qout, cnt_out);
        }
    }
}

// ---------------------------------------------------------------------
// Load-balanced expansion (SCC_EXPAND_LB).  A 256-lane workgroup takes 256
// frontier vertices (or, when tiles are fewer than workgroups, one slice of
// their arcs), scans their degrees in LDS and walks the concatenated
// adjacency slots with all lanes, so consecutive lanes read consecutive
// column entries whatever the degree mix.  A vertex of degree >= LB_BIG
// (a hub) is instead appended to this level's hub list; the next level's
// launch first splits the listed hubs into LB_CHUNK-arc chunks spread over
// every workgroup, then takes its own tiles.  A hub's neighbours thus join
// the frontier one level late, which leaves every reachable set unchanged
// and makes a level a single launch.
constexpr uint32_t LB_TILE = 256;
constexpr uint32_t LB_CHUNK = 512;
constexpr uint32_t LB_SLICES = 16;   // most workgroups sharing one tile
constexpr uint32_t LB_DIRECT_N = 8192;   // frontier and mean degree of claims
constexpr uint32_t LB_DIRECT_DEG = 4;    // without a visited-word read first
constexpr uint32_t LB_HOPS = 64;         // local continuation: hops per workgroup and launch
constexpr uint32_t HOP_LANE_DEG = 4;     // a continuation hop runs a vertex per lane up to this degree
static_assert(LB_TILE * HOP_LANE_DEG <= QBUF, "a lane hop's claims fit the LDS queue");

// Exclusive scan of x over the 256-lane workgroup; returns the prefix and
// writes the total to *total.  `wsum` is LDS scratch of 4 words.
__device__ __forceinline__ uint32_t block_exscan256(uint32_t x, uint32_t *wsum, uint32_t *total)
{
    const uint32_t l = wlane(), w = threadIdx.x >> 6;
    uint32_t incl = x;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(incl, o);
        if (l >= (uint32_t)o) incl += y;
    }
    if (l == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t base = 0, tot = 0;
    for (uint32_t i = 0; i < 4; i++) {
        uint32_t s = wsum[i];
        base += i < w ? s : 0u;
        tot += s;
    }
    *total = tot;
    return base + incl - x;
}

// LDS words of a top-down level: the staging queue and its two words, then
// the tile's adjacency offsets, row starts, classes and scan scratch.
constexpr uint32_t LB_LDS = QBUF + 2 + (LB_TILE + 1) + 3 * LB_TILE + 4;

// One top-down level of direction D for workgroup `bid` of `bpd`; `lds`
// holds LB_LDS words.  The input queue holds n entries: from index 0, or,
// after a bottom-up level (pre not null), in segments read through seg_pos
// with the prefixes `pre` and slot regions of qcap.  Called by the whole
// workgroup.
template <bool SINGLE, bool COUNT = false>
__device__ __forceinline__ void lb_level(const DirArgs &D, const uint32_t *cls, int L, uint32_t bid,
                                         uint32_t bpd, uint32_t *lds, uint32_t n, const uint32_t *pre,
                                         uint32_t qcap, WorkCount &wc)
{
    const unsigned long long packed = *D.pbigc;
    const uint32_t items = (uint32_t)(packed >> 32), chunks = (uint32_t)packed;
    uint32_t *cnt_out = &D.cnt[CT((L + 1) % 3)];
    if (bid == 0) {
        seg_clear(&D.cnt[CT((L + 2) % 3)], &D.cslot[CT(((L + 2) % 3) * kSpread)]);
        if (threadIdx.x == 0) *D.bigc_zero = 0ull;
    }
    if (n == 0 && items == 0) return;
    const LdsQueue q{lds, lds + QBUF, lds + QBUF + 1};
    uint32_t *s_off = lds + QBUF + 2;
    uint32_t *s_beg = s_off + LB_TILE + 1;
    uint32_t *s_cls = s_beg + LB_TILE;
    uint32_t *s_wsum = s_cls + LB_TILE;
    if (threadIdx.x == 0) *q.n = 0;
    __syncthreads();
    const uint32_t *__restrict__ R = D.R;
    const uint32_t *__restrict__ C = D.C;
    // hubs of the previous level: chunk g belongs to the item whose chunk
    // offset is the last one <= g (offsets grow with the item index because
    // both come from one 64-bit atomicAdd)
    for (uint32_t gch = bid; gch < chunks; gch += bpd) {
        uint32_t lo = 0, hi = items - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (D.pbigoff[mid] <= gch) lo = mid;
            else hi = mid - 1;
        }
        const uint32_t v = D.pbig[lo];
        const uint32_t c = SINGLE ? 0u : cls[v];
        const uint32_t b = R[v] + (gch - D.pbigoff[lo]) * LB_CHUNK;
        const uint32_t e = min(R[v + 1], b + LB_CHUNK);
        for (uint32_t j0 = b; j0 < e; j0 += 256) {
            const uint32_t j = j0 + threadIdx.x;
            bool claimed = false;
            uint32_t w = 0;
            if (j < e) {
                w = C[j];
                claimed = try_claim<SINGLE>(D.vis, cls, w, c);
                if (COUNT) {
                    wc.n[W_TDE]++;
                    wc.n[W_CLAIM] += claimed;
                }
            }
            lq_push(q, claimed, w);
            lq_flush(q, cnt_out, D.qout, false);
        }
    }
    // With fewer tiles than workgroups, each tile's arcs are cut into
    // `slices` equal ranges taken by different workgroups (a few hundred
    // frontier vertices of degree up to LB_BIG would otherwise keep a few
    // workgroups busy for the whole level while the rest idle).  Every
    // slice scans the tile's degrees; slice 0 alone lists its hubs.
    const uint32_t tiles = (n + LB_TILE - 1) / LB_TILE;
    const uint32_t slices = (tiles == 0 || tiles >= bpd) ? 1u : min(LB_SLICES, bpd / tiles);
    for (uint32_t unit = bid; unit < tiles * slices; unit += bpd) {
        const uint32_t base = (unit / slices) * LB_TILE, slice = unit % slices;
        const uint32_t i = base + threadIdx.x;
        uint32_t v = 0, beg = 0, deg = 0, c = 0;
        if (i < n) {
            v = D.qin[pre ? seg_pos(pre, i, qcap) : i];
            beg = R[v];
            deg = R[v + 1] - beg;
            if (!SINGLE) c = cls[v];
            if (COUNT && slice == 0) wc.n[W_TDV]++;
            if (deg >= LB_BIG && slice == 0) {
                const unsigned long long nch = (deg + LB_CHUNK - 1) / LB_CHUNK;
                const unsigned long long old = atomicAdd(D.bigc, (1ull << 32) | nch);
                const uint32_t item = (uint32_t)(old >> 32);
                D.big[item] = v;
                D.bigoff[item] = (uint32_t)old;
            }
            if (deg >= LB_BIG) deg = 0;
        }
        uint32_t total;
        const uint32_t off = block_exscan256(deg, s_wsum, &total);
        s_off[threadIdx.x] = off;
        s_beg[threadIdx.x] = beg;
        s_cls[threadIdx.x] = c;
        if (threadIdx.x == 0) s_off[LB_TILE] = total;
        __syncthreads();
        const uint32_t jlo = (uint32_t)((uint64_t)total * slice / slices);
        const uint32_t jhi = (uint32_t)((uint64_t)total * (slice + 1) / slices);
        // a small, low-degree frontier of a single-class search (the long
        // tail of a high-diameter graph) claims without reading the visited
        // word first: one dependent load less per level, and at most
        // LB_DIRECT_N * LB_DIRECT_DEG atomics per level
        const bool direct = SINGLE && n <= LB_DIRECT_N && total <= LB_DIRECT_DEG * LB_TILE;
        for (uint32_t j0 = jlo; j0 < jhi; j0 += LB_TILE) {
            const uint32_t j = j0 + threadIdx.x;
            bool claimed = false;
            uint32_t w = 0;
            if (j < jhi) {
                // owner = last k with s_off[k] <= j (degree-0 entries share offsets)
                uint32_t lo = 0, hi = LB_TILE - 1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (s_off[mid] <= j) lo = mid;
                    else hi = mid - 1;
                }
                w = C[s_beg[lo] + (j - s_off[lo])];
                const uint32_t bit = 1u << (w & 31);
                claimed = direct ? !(atomicOr(&D.vis[vword(w)], bit) & bit)
                                 : try_claim<SINGLE>(D.vis, cls, w, SINGLE ? 0u : s_cls[lo]);
                if (COUNT) {
                    wc.n[W_TDE]++;
                    wc.n[W_CLAIM] += claimed;
                }
            }
            lq_push(q, claimed, w);
            lq_flush(q, cnt_out, D.qout, false);
        }
        __syncthreads();   // LDS reuse by the next tile
    }
    // Local continuation.  In a narrow single-class level (a high-diameter
    // graph's search) each workgroup goes on expanding the vertices it has
    // just claimed, for up to LB_HOPS more hops, while they fit one tile and
    // carry at most LB_DIRECT_DEG * LB_TILE arcs; what it claims in the last
    // hop, or when it stops early, goes to the next level's queue as usual.
    // Reachability does not depend on the order of expansion, and each
    // vertex is still claimed once, so the reachable sets are unchanged;
    // the search needs far fewer launches (each a chain of global round
    // trips) when the frontier stays narrow for thousands of levels.
    //
    // A hop whose vertices all have at most HOP_LANE_DEG arcs (a grid, a
    // sparse tail) runs one vertex per lane: no scan, and each claim's row
    // offsets are loaded beside its claim atomic and kept in LDS with the
    // queue entry, so a hop is two dependent global round trips (column,
    // then claim and rows) instead of three.
    if (SINGLE && n <= LB_DIRECT_N) {
        uint32_t *s_hb = s_off, *s_he = s_beg;   // rows of queue entries < LB_TILE
        uint32_t *const vis = D.vis;
        bool rows = false;                        // entries carry their rows
        for (uint32_t hop = 0; hop < LB_HOPS; hop++) {
            __syncthreads();
            const uint32_t m = *q.n;
            if (m == 0 || m > LB_TILE) break;
            const uint32_t t = threadIdx.x;
            const uint32_t v = t < m ? q.buf[t] : 0u;
            uint32_t beg = 0, end = 0;
            if (t < m) {
                beg = rows ? s_hb[t] : R[v];
                end = rows ? s_he[t] : R[v + 1];
            }
            const uint32_t deg = end - beg;
            // every lane has read its entry and rows once this barrier passes
            if (!__syncthreads_or(deg > HOP_LANE_DEG)) {
                if (t == 0) *q.n = 0;
                if (COUNT) {
                    wc.n[W_TDV] += t < m;
                    wc.n[W_TDE] += deg;
                }
                __syncthreads();
                // Straight-line per lane with arcs: every load and claim is
                // issued before any result is waited for (a claim per branch
                // would wait for the previous one).  Slots past the degree
                // repeat the last arc with an empty mask, which changes
                // nothing and claims nothing.
                uint32_t w[HOP_LANE_DEG] = {}, rb[HOP_LANE_DEG] = {}, re[HOP_LANE_DEG] = {};
                uint32_t old[HOP_LANE_DEG] = {};
                bool claimed[HOP_LANE_DEG];
                if (deg > 0) {
#pragma unroll
                    for (uint32_t k = 0; k < HOP_LANE_DEG; k++) w[k] = C[beg + min(k, deg - 1)];
#pragma unroll
                    for (uint32_t k = 0; k < HOP_LANE_DEG; k++)
                        old[k] = atomicOr(&vis[vword(w[k])], k < deg ? 1u << (w[k] & 31) : 0u);
#pragma unroll
                    for (uint32_t k = 0; k < HOP_LANE_DEG; k++) {
                        rb[k] = R[w[k]];
                        re[k] = R[w[k] + 1];
                    }
                }
#pragma unroll
                for (uint32_t k = 0; k < HOP_LANE_DEG; k++)
                    claimed[k] = k < deg && !(old[k] & (1u << (w[k] & 31)));
#pragma unroll
                for (uint32_t k = 0; k < HOP_LANE_DEG; k++) {
                    if (COUNT) wc.n[W_CLAIM] += claimed[k];
                    const uint64_t act = __ballot(1), mk = __ballot(claimed[k]);
                    if (!mk) continue;
                    const int leader = __ffsll((unsigned long long)act) - 1;
                    uint32_t base = 0;
                    if ((int)wlane() == leader) base = atomicAdd(q.n, (uint32_t)__popcll(mk));
                    base = __shfl(base, leader);
                    if (claimed[k]) {
                        const uint32_t pos = base + (uint32_t)__popcll(mk & ((1ull << wlane()) - 1));
                        q.buf[pos] = w[k];   // pos < LB_TILE * HOP_LANE_DEG <= QBUF
                        if (pos < LB_TILE) {
                            s_hb[pos] = rb[k];
                            s_he[pos] = re[k];
                        }
                    }
                }
                rows = true;
                continue;
            }
            uint32_t total;
            const uint32_t off = block_exscan256(deg, s_wsum, &total);
            if (total > LB_DIRECT_DEG * LB_TILE) break;   // uniform: the queue is untouched
            s_off[t] = off;
            s_beg[t] = beg;
            if (t == 0) s_off[LB_TILE] = total;
            __syncthreads();   // every lane holds its entry before the queue is reused
            if (t == 0) *q.n = 0;
            if (COUNT) wc.n[W_TDV] += t < m;
            __syncthreads();
            for (uint32_t j0 = 0; j0 < total; j0 += LB_TILE) {
                const uint32_t j = j0 + t;
                bool claimed = false;
                uint32_t w = 0;
                if (j < total) {
                    uint32_t lo = 0, hi = LB_TILE - 1;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi + 1) >> 1;
                        if (s_off[mid] <= j) lo = mid;
                        else hi = mid - 1;
                    }
                    w = C[s_beg[lo] + (j - s_off[lo])];
                    const uint32_t bit = 1u << (w & 31);
                    claimed = !(atomicOr(&D.vis[vword(w)], bit) & bit);
                    if (COUNT) {
                        wc.n[W_TDE]++;
                        wc.n[W_CLAIM] += claimed;
                    }
                }
                lq_push(q, claimed, w);
                lq_flush(q, cnt_out, D.qout, false);
            }
            rows = false;
        }
    }
    lq_flush(q, cnt_out, D.qout, true);
}

template <bool SINGLE>
__global__ __launch_bounds__(256) void expand_lb_kernel(LevelArgs a, uint32_t bpd)
{
    __shared__ uint32_t s_lds[LB_LDS];
    int dir;
    uint32_t bid;
    split_dirs(a.ndirs, bpd, dir, bid);
    const DirArgs &D = a.d[dir];
    // no bottom-up level runs in this mapping: the queues start at index 0
    WorkCount wc;
    lb_level<SINGLE>(D, a.cls, a.level, bid, bpd, s_lds, D.cnt[CT(a.level % 3)], nullptr, 0u, wc);
}

static void launch_lb(hipStream_t s, const LevelArgs &a, bool single, uint32_t bpd)
{
    bpd = dir_blocks(bpd, a.ndirs);
    const uint32_t grid = bpd * (uint32_t)a.ndirs;
    if (single) expand_lb_kernel<true><<<grid, 256, 0, s>>>(a, bpd);
    else expand_lb_kernel<false><<<grid, 256, 0, s>>>(a, bpd);
}

// ---------------------------------------------------------------------
// Bottom-up pass (direction-optimising reachability, Beamer et al. SC'12):
// a live vertex u not yet visited scans its parents and joins the search as
// soon as one of them is visited and in u's class.  Only u's own lane sets
// u's bit, so there is one atomic per newly visited vertex instead of one
// per probed arc.  A parent that became visited during this same pass may
// be missed (stale cache line); the next pass sees it because kernel
// boundaries make every write visible.  The pass pushes the vertices it
// visits, so a following top-down level can continue from them.
//
// Two tiers: each lane probes the first parent of its vertex, read from
// the first-parent array (one coalesced word per vertex, loaded with the
// row offsets); vertices with longer parent lists and no hit yet are listed
// in LDS and scanned by groups of BU_GROUP lanes, BU_GROUP parents per
// step, with a ballot to stop each group at its first hit.  One parent up
// front and groups of four were fastest on R-MAT scales 22 and 24 (probing
// more parents before knowing whether the first one hits reads lines that
// are mostly not needed).
//
// The pass is wave-centric: each wave streams its own 64-vertex chunks,
// two at a time (each lane carries one vertex of each, so twice as many
// chains of dependent loads -- vertex word, row offsets, parent ids, parent
// bits -- are in flight), with a private LDS queue and long list, and never
// waits on its workgroup.  A wave whose queue fills reserves global space
// itself; what remains at the end goes out with one atomic per workgroup
// (same-address atomics serialise).
//
// Claims need no atomics: chunks start at multiples of 64, so the two
// visited words of a chunk belong to one wave for the whole pass, and the
// wave ORs its hits into each with one plain store (device-wide atomics
// run at about 10 G/s on MI355X, which would bound a pass that visits
// millions of vertices).
constexpr uint32_t BU_SHORT = 1;   // the first parent, read from the first-parent array
constexpr uint32_t BU_GROUP = 4;
constexpr uint32_t BU_WQ = 1024;
constexpr uint32_t BU_CHUNKS = 2;   // 64-vertex chunks per wave iteration
constexpr uint32_t BU_SPARSE = 16;  // unvisited live vertices under V / BU_SPARSE: a sparse level
// 64-vertex chunks a sparse step scans.  Its candidate masks and counts are
// wave-uniform and live in scalar registers, which every level of the
// kernel then reserves; expand_do_kernel is held to 80 of them.
constexpr uint32_t BU_SP_CHUNKS = 6;
constexpr uint32_t BU_WH = (BU_SP_CHUNKS > BU_CHUNKS ? BU_SP_CHUNKS : BU_CHUNKS) * 2;   // hit words per wave

// Position of the r-th set bit of m (r < popcount(m)), by halving.
__device__ __forceinline__ uint32_t nth_bit(uint64_t m, uint32_t r)
{
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t w = 32; w >= 1; w >>= 1) {
        const uint32_t c = (uint32_t)__popcll(m & ((1ull << w) - 1));
        const bool skip = r >= c;
        r -= skip ? c : 0u;
        m = skip ? m >> w : m;
        pos += skip ? w : 0u;
    }
    return pos;
}

// One 64-bit load gives p's visited-or-finished and finished bits.
template <bool SINGLE>
__device__ __forceinline__ bool visited_parent(const uint32_t *vis, const uint32_t *cls, uint32_t p, uint32_t c)
{
    const uint64_t w = reinterpret_cast<const uint64_t *>(vis)[p >> 5];
    const uint32_t pb = 1u << (p & 31);
    return ((uint32_t)w & ~(uint32_t)(w >> 32) & pb) && (SINGLE || cls[p] == c);
}

// LDS words of a bottom-up level: each wave's queue, long list and the
// hits of its chunks (two words per chunk) that lanes other than the
// vertex's own found.
constexpr uint32_t BU_LDS = 4 * BU_WQ + 4 * BU_CHUNKS * 64 + 4 * BU_WH;

// One bottom-up level of direction D (parents in PR/PC) for workgroup `bid`
// of `bpd`; `lds` holds BU_LDS words.  Called by the whole workgroup.
//
// A sparse level (few unvisited vertices left) has one or two candidates
// per 64-vertex chunk, and sweeping it a chunk per lane set ran the whole
// chunk body for them: the level was bound by instructions, not memory
// (43 us for 0.2M L2 requests on R-MAT scale 24).  A sparse step instead
// reads the visited words of BU_SP_CHUNKS of the wave's chunks, lists their
// candidates across the lanes (candidate i of the step on lane i % 64, its
// chunk from the prefix of the chunks' candidate counts and its position
// from nth_bit) and probes them 64 at a time, one per lane; hits are ORed
// into per-chunk masks in LDS and stored by the lanes that own the chunks'
// visited words, as in a dense step.
//
// The vertices a pass visits go to the region of slot bid % kSpread of the
// output queue (slot_cap(V, bpd) = qcap entries: the slot's workgroups own
// that many vertices at most) and are counted in that slot of the count, so
// the flushes of the pass's workgroups are spread over kSpread addresses.
template <bool SINGLE, bool COUNT = false>
__device__ __forceinline__ void bu_level(const DirArgs &D, const uint32_t *__restrict__ PR,
                                         const uint32_t *__restrict__ PC, const uint32_t *__restrict__ PF,
                                         const uint32_t *cls, uint32_t V, int L, uint32_t bid, uint32_t bpd,
                                         uint32_t *lds, uint32_t qcap, uint32_t *anom, bool sparse, WorkCount &wc)
{
    const uint32_t slot = bid % kSpread;
    uint32_t *cnt_out = &D.cslot[CT(((L + 1) % 3) * kSpread + slot)];
    uint32_t *qout = D.qout + (size_t)slot * qcap;
    if (bid == 0) {
        seg_clear(&D.cnt[CT((L + 2) % 3)], &D.cslot[CT(((L + 2) % 3) * kSpread)]);
        if (threadIdx.x == 0) *D.bigc_zero = 0ull;
    }
    const uint32_t wv = threadIdx.x >> 6, lane = wlane();
    uint32_t *wq = lds + wv * BU_WQ;
    uint32_t *wl = lds + 4 * BU_WQ + wv * BU_CHUNKS * 64;
    uint32_t *wh = lds + 4 * BU_WQ + 4 * BU_CHUNKS * 64 + wv * BU_WH;
    uint32_t nq = 0;
    const uint32_t sub = lane % BU_GROUP, grp = lane / BU_GROUP;
    const uint32_t gshift = lane & ~(BU_GROUP - 1);
    const uint32_t stride = bpd * 256;   // one chunk of every wave of the grid
    const uint64_t *vis64 = reinterpret_cast<const uint64_t *>(D.vis);
    // long parent lists of wl[0, nlong): 64 / BU_GROUP vertices per round,
    // each group stops at its first hit; a hit goes to its chunk's mask in wh
    auto long_lists = [&](uint32_t base, uint32_t nlong) {
        for (uint32_t k0 = 0; k0 < nlong; k0 += 64 / BU_GROUP) {
            const uint32_t k = k0 + grp;
            bool hit = false;
            uint32_t x = 0;
            if (k < nlong) {
                x = wl[k];
                const uint32_t cx = SINGLE ? 0u : cls[x];
                const uint32_t ex = PR[x + 1];
                for (uint32_t j0 = PR[x] + BU_SHORT; j0 < ex && !hit; j0 += BU_GROUP) {
                    const uint32_t j = j0 + sub;
                    const bool h = j < ex && visited_parent<SINGLE>(D.vis, cls, PC[j], cx);
                    if (COUNT) wc.n[W_BUM] += j < ex;
                    hit = ((__ballot(h) >> gshift) & ((1ull << BU_GROUP) - 1)) != 0;
                }
            }
            const bool claimed = hit && sub == 0;
            if (COUNT) wc.n[W_CLAIM] += claimed;
            if (claimed) {
                const uint32_t hx = (x - base) / stride, off = x - base - hx * stride;
                atomicOr(&wh[hx * 2 + (off >> 5)], 1u << (off & 31));   // LDS
            }
            wave_push(wq, nq, claimed, x);
        }
    };
    if (!sparse) {
        for (uint32_t base = (bid * 4 + wv) * 64; base <= V; base += BU_CHUNKS * stride) {
            if (nq + 2 * BU_CHUNKS * 64 > BU_WQ) wave_flush(wq, nq, cnt_out, qout, qcap, anom);
            uint32_t u[BU_CHUNKS], word[BU_CHUNKS], c[BU_CHUNKS], b[BU_CHUNKS], e[BU_CHUNKS];
            bool cand[BU_CHUNKS], found[BU_CHUNKS], need_long[BU_CHUNKS];
            uint64_t hits[BU_CHUNKS];
#pragma unroll
            for (uint32_t h = 0; h < BU_CHUNKS; h++) {
                // lane l of chunk h holds vertex base + h * stride + l (vertex 0,
                // the dummy, is finished)
                u[h] = base + h * stride + lane;
                word[h] = u[h] <= V ? D.vis[vword(u[h])] : ~0u;
                cand[h] = u[h] != 0 && !(word[h] & (1u << (u[h] & 31)));
            }
            // the first parent comes from PF (coalesced; 0: no parent), not from
            // a random line of the parent list, and only a candidate whose first
            // parent misses loads its row offsets to see whether more follow.
            // PF is read for every lane beside its visited word (past V: vertex
            // 0's entry) rather than after it for candidates only: one
            // dependent step less per chunk for a coalesced word per vertex
            // (R-MAT scale 24's search 0.382 -> 0.375 ms)
            uint32_t pf[BU_CHUNKS];
#pragma unroll
            for (uint32_t h = 0; h < BU_CHUNKS; h++) pf[h] = PF[u[h] <= V ? u[h] : 0u];
#pragma unroll
            for (uint32_t h = 0; h < BU_CHUNKS; h++) {
                c[h] = (SINGLE || !cand[h]) ? 0u : cls[u[h]];
                if (!cand[h]) pf[h] = 0u;
            }
            // Every chunk's parent probe is issued before any result is used,
            // and then every chunk's row offsets, without per-lane branches: a
            // lane with nothing to load reads vertex 0's entries (the dummy:
            // group 0 of the bitmap, rows 0 and 1), which are always in bounds.
            // A wave with nothing to probe (or no row to read) skips the loads.
            // With a probe and its rows per chunk in one branch, each load
            // waited for the one before it.
            bool anyp = false;
#pragma unroll
            for (uint32_t h = 0; h < BU_CHUNKS; h++) anyp |= pf[h] != 0;
            uint64_t pw[BU_CHUNKS] = {};
            if (__ballot(anyp)) {
#pragma unroll
                for (uint32_t h = 0; h < BU_CHUNKS; h++) pw[h] = reinterpret_cast<const uint64_t *>(D.vis)[pf[h] >> 5];
            }
            bool more[BU_CHUNKS];
#pragma unroll
            for (uint32_t h = 0; h < BU_CHUNKS; h++) {
                // visited in this search (visited-or-finished and not finished),
                // and of the candidate's class
                const bool vp = ((uint32_t)pw[h] & ~(uint32_t)(pw[h] >> 32)) & (1u << (pf[h] & 31));
                found[h] = pf[h] && vp && (SINGLE || cls[pf[h]] == c[h]);
                more[h] = pf[h] && !found[h];
            }
            bool anym = false;
#pragma unroll
            for (uint32_t h = 0; h < BU_CHUNKS; h++) {
                b[h] = e[h] = 0;
                anym |= more[h];
            }
            if (__ballot(anym)) {
#pragma unroll
                for (uint32_t h = 0; h < BU_CHUNKS; h++) {
                    const uint32_t r = more[h] ? u[h] : 0u;
                    b[h] = PR[r];
                    e[h] = PR[r + 1];
                }
#pragma unroll
                for (uint32_t h = 0; h < BU_CHUNKS; h++) {
                    if (!more[h]) b[h] = e[h] = 0;
                }
            }
#pragma unroll
            for (uint32_t h = 0; h < BU_CHUNKS; h++) {
                hits[h] = __ballot(found[h]);
                need_long[h] = more[h] && e[h] - b[h] > BU_SHORT;
                if (COUNT) {
                    // a vertex whose visited bit and first-parent word were read
                    wc.n[W_BUS] += u[h] <= V;
                    wc.n[W_BUC] += cand[h];
                    wc.n[W_BUF] += pf[h] != 0;
                    wc.n[W_BUR] += more[h];
                    wc.n[W_CLAIM] += found[h];
                }
            }
            uint32_t nlong = 0;
#pragma unroll
            for (uint32_t h = 0; h < BU_CHUNKS; h++) {
                wave_push(wq, nq, found[h], u[h]);
                const uint64_t lm = __ballot(need_long[h]);
                if (need_long[h]) wl[nlong + __popcll(lm & ((1ull << lane) - 1))] = u[h];
                nlong += (uint32_t)__popcll(lm);
            }
            if (lane < BU_CHUNKS * 2) wh[lane] = 0;
            wave_sync();
            long_lists(base, nlong);
            wave_sync();   // long-list hits are in wh; the next iteration rewrites wl
#pragma unroll
            for (uint32_t h = 0; h < BU_CHUNKS; h++) {
                // lanes 0 and 32 own the chunk's two visited words
                const uint32_t m = (lane == 0 ? (uint32_t)hits[h] : (uint32_t)(hits[h] >> 32)) | wh[h * 2 + (lane >> 5)];
                if ((lane & 31) == 0 && m) D.vis[vword(u[h])] = word[h] | m;
            }
        }
    } else {
        for (uint32_t base = (bid * 4 + wv) * 64; base <= V; base += BU_SP_CHUNKS * stride) {
            // a step pushes each of its candidates at most once
            if (nq + BU_SP_CHUNKS * 64 > BU_WQ) wave_flush(wq, nq, cnt_out, qout, qcap, anom);
            uint32_t word[BU_SP_CHUNKS];
#pragma unroll
            for (uint32_t h = 0; h < BU_SP_CHUNKS; h++) {
                const uint32_t uh = base + h * stride + lane;
                word[h] = uh <= V ? D.vis[vword(uh)] : ~0u;
            }
            uint64_t cm[BU_SP_CHUNKS];
            uint32_t pre[BU_SP_CHUNKS + 1];
            pre[0] = 0;
#pragma unroll
            for (uint32_t h = 0; h < BU_SP_CHUNKS; h++) {
                const uint32_t uh = base + h * stride + lane;   // vertex 0, the dummy, is finished
                cm[h] = __ballot(uh != 0 && !(word[h] & (1u << (uh & 31))));
                pre[h + 1] = pre[h] + (uint32_t)__popcll(cm[h]);
            }
            const uint32_t total = pre[BU_SP_CHUNKS];
            if (!total) continue;
            if (lane < BU_SP_CHUNKS * 2) wh[lane] = 0;
            wave_sync();
            for (uint32_t b0 = 0; b0 < total; b0 += 64) {
                const uint32_t i = b0 + lane;
                const bool ok = i < total;
                // chunk h holds candidates [pre[h], pre[h + 1]) of the step
                uint32_t h = 0;
#pragma unroll
                for (uint32_t k = 1; k < BU_SP_CHUNKS; k++) h += i >= pre[k];
                uint64_t m = cm[0];
                uint32_t r0 = 0;
#pragma unroll
                for (uint32_t k = 1; k < BU_SP_CHUNKS; k++) {
                    m = h == k ? cm[k] : m;
                    r0 = h == k ? pre[k] : r0;
                }
                const uint32_t pos = ok ? nth_bit(m, i - r0) : 0u;
                const uint32_t u = ok ? base + h * stride + pos : 0u;   // an idle lane reads vertex 0's words
                // the first parent and the class together, then the parent's
                // bitmap pair, then (after a miss) the row offsets
                uint32_t pf = PF[u];
                const uint32_t c = SINGLE ? 0u : cls[u];
                if (!ok) pf = 0u;
                uint64_t pw = 0;
                if (__ballot(pf != 0)) pw = vis64[pf >> 5];
                const bool vp = ((uint32_t)pw & ~(uint32_t)(pw >> 32)) & (1u << (pf & 31));
                const bool found = pf && vp && (SINGLE || cls[pf] == c);
                const bool more = pf && !found;
                uint32_t rb = 0, re = 0;
                if (__ballot(more)) {
                    const uint32_t r = more ? u : 0u;
                    rb = PR[r];
                    re = PR[r + 1];
                    if (!more) rb = re = 0;
                }
                const bool need_long = more && re - rb > BU_SHORT;
                if (COUNT) {
                    wc.n[W_BUS] += ok;   // a candidate of a sparse step read its bit and first parent
                    wc.n[W_BUC] += ok;
                    wc.n[W_BUF] += pf != 0;
                    wc.n[W_BUR] += more;
                    wc.n[W_CLAIM] += found;
                }
                if (found) atomicOr(&wh[h * 2 + (pos >> 5)], 1u << (pos & 31));   // LDS
                wave_push(wq, nq, found, u);
                const uint64_t lm = __ballot(need_long);
                if (need_long) wl[__popcll(lm & ((1ull << lane) - 1))] = u;
                wave_sync();
                long_lists(base, (uint32_t)__popcll(lm));
                wave_sync();   // the hits are in wh; the next batch rewrites wl
            }
#pragma unroll
            for (uint32_t h = 0; h < BU_SP_CHUNKS; h++) {
                // lanes 0 and 32 own the chunk's two visited words
                const uint32_t mh = wh[h * 2 + (lane >> 5)];
                if ((lane & 31) == 0 && mh) D.vis[vword(base + h * stride + lane)] = word[h] | mh;
            }
            wave_sync();   // the next step clears wh
        }
    }
    wave_flush_block(wq, nq, cnt_out, qout, qcap, anom);
}

// ---------------------------------------------------------------------
// Direction-optimised level decided on the device.  Every workgroup of the
// level-L launch recomputes each direction's mode from the mode and visited
// count the level L-1 launch stored (slot (L-1)&1) and the counters level
// L-1 left (vertices queued, hubs deferred); workgroup 0 of each direction
// stores the result in slot L&1 for level L+1.  So the host can queue any
// number of levels without reading counters in between; a direction that
// has finished turns its workgroups into no-ops.
// Adds each lane's work units to the counters, one 64-bit atomic per wave
// and unit (the counted run only).
__device__ __forceinline__ void add_work(uint32_t *ctr, const WorkCount &wc)
{
#pragma unroll
    for (int k = 0; k < 8; k++) {
        uint32_t x = wc.n[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
        if (wlane() == 0 && x)
            atomicAdd(reinterpret_cast<unsigned long long *>(ctr + CT(C_WORK + k)), (unsigned long long)x);
    }
}

template <bool SINGLE, bool COUNT>
// At most 80 scalar registers (the compiler moves the rest to vector lanes):
// with 86 to 102 (the sparse steps' uniform masks) R-MAT scale 24's dense
// bottom-up level took 251 us instead of 225, although the compiler
// reported 8 waves per SIMD for 86 and 94; held to 80, the same code runs
// the dense level in 224 us and the two sparse levels in 52 + 23 us.
__global__ __launch_bounds__(256) __attribute__((amdgpu_num_sgpr(80))) void expand_do_kernel(DoArgs a, uint32_t bpd)
{
    __shared__ uint32_t s_lds[BU_LDS > LB_LDS ? BU_LDS : LB_LDS];
    __shared__ uint32_t s_pre[kSpread + 2];
    int dir;
    uint32_t bid;
    split_dirs(a.ndirs, bpd, dir, bid);
    const DirArgs &D = a.d[dir];
    const int L = a.level;
    DirState st = load_dir_state(a.ctr, L == 0 ? 0 : (L - 1) & 1, dir);
    // the input queue: after a bottom-up level (the mode level L-1 ran in)
    // its count is the own word plus the slots, one load per lane of wave 0;
    // after a top-down level or the seeds the own word alone, with no load
    // added to the level's first dependent step
    const bool slotted = L > 0 && st.mode == DIR_BU;
    uint32_t n;
    if (slotted) {
        seg_prefix(&D.cnt[CT(L % 3)], &D.cslot[CT((L % 3) * kSpread)], a.qcap, s_pre);
        n = s_pre[kSpread + 1];
    } else {
        n = D.cnt[CT(L % 3)];
    }
    if (L > 0) {
        st = dir_decide(st, n, (uint32_t)(*D.pbigc >> 32), a.ctr[CT(C_NLIVE)], (uint32_t)L, a.bu_ok != 0);
        if (bid == 0 && threadIdx.x == 0) store_dir_state(a.ctr, L & 1, dir, st);
    }
    WorkCount wc;
    if (st.mode == DIR_TD)
        lb_level<SINGLE, COUNT>(D, a.cls, L, bid, bpd, s_lds, n, slotted ? s_pre : nullptr, a.qcap, wc);
    else if (st.mode == DIR_BU)
        bu_level<SINGLE, COUNT>(D, a.PR[dir], a.PC[dir], a.PF[dir], a.cls, a.V, L, bid, bpd, s_lds, a.qcap,
                                a.ctr + CT(C_ANOM),
                                (int64_t)a.ctr[CT(C_NLIVE)] - (int64_t)st.visited < (int64_t)(a.V / BU_SPARSE), wc);
    if (COUNT) add_work(a.ctr, wc);
}

// ---------------------------------------------------------------------
// Narrow frontiers: many top-down levels inside one workgroup per direction.
// A high-diameter graph (the 4096 x 4096 grid: 8,193 levels, frontiers of
// a few thousand vertices) pays a launch and a chain of global round trips
// (queue count, queue, row offsets, columns, claim, queue reservation) per
// device-wide level.  Here one 1,024-lane workgroup per direction keeps the
// frontier in LDS and loops over levels with workgroup barriers only: per
// level the row offsets of the frontier, an LDS scan of the degrees, the
// columns and the claims (workgroup-scope atomics: during this kernel no
// other workgroup touches this direction's bitmap), and an LDS append.
//
// The kernel takes over at level L0 in the state the level-L0 launch of
// expand_do_kernel would start from and leaves that same position rewritten
// for the device-wide levels that follow: the frontier it stops at is level
// L0's input queue and count, and the search state of slot (L0-1)&1 (slot
// 0 for L0 = 0) is the one the last expanded level would have left.  It
// stops when the frontier is empty, when the next frontier exceeds NL_SPILL
// (claims past NL_CAP go straight to the global queue), or when a level
// after the first would scan more than NL_ARCS arcs.  It does nothing
// unless level L0 would run top-down from a top-down level's queue with no
// hubs pending and at most NL_CAP vertices, so the host may queue it behind
// levels whose outcome it has not seen.  At least one level is expanded
// whenever it runs, and every level either claims a vertex or ends the
// loop, so it ends.
//
// One CU's scattered loads and claims limit a level to about 12 ns per
// frontier vertex on top of a 2.6 us chain (measured: a 20,000-vertex
// cycle, one vertex per level, 2.6 us per level against 6.2 us device-wide;
// the 4096^2 grid's frontiers of ~2,000 vertices, 27 us against 7.2 us), so
// the host starts it on frontiers of at most NL_ENTER vertices and it hands
// back beyond NL_SPILL.  The host uses it for multi-class searches only: a
// single-class search's top-down levels continue narrow frontiers inside
// the device-wide launch (lb_level's local continuation), which is faster.
constexpr uint32_t NL_THREADS = 1024;
constexpr uint32_t NL_CAP = 6144;                 // frontier entries held in LDS
constexpr uint32_t NL_ENTER = 256;                // host: start on frontiers this small
constexpr uint32_t NL_SPILL = 512;                // hand back when the next frontier is larger
constexpr uint32_t NL_PER = NL_CAP / NL_THREADS;  // frontier entries per lane
constexpr uint32_t NL_ARCS = 32768;               // larger levels go back to the device-wide kernel
constexpr uint32_t NL_K = 8;                      // arcs per lane per round

// Exclusive scan of deg[0..n) in place over the workgroup (n <= NL_CAP):
// each lane sums its NL_PER consecutive entries, a wave scan and a scan of
// the 16 wave totals place them.  Returns the total.
__device__ __forceinline__ uint32_t nl_scan(uint32_t *deg, uint32_t n, uint32_t *s_wsum)
{
    const uint32_t t = threadIdx.x, l = wlane(), w = t >> 6;
    uint32_t loc[NL_PER], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < NL_PER; k++) {
        const uint32_t i = t * NL_PER + k;
        loc[k] = i < n ? deg[i] : 0u;
        sum += loc[k];
    }
    uint32_t incl = sum;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o);
        if (l >= (uint32_t)o) incl += y;
    }
    if (l == 63) s_wsum[w] = incl;
    __syncthreads();
    uint32_t base = 0, total = 0;
    for (uint32_t i = 0; i < NL_THREADS / 64; i++) {
        const uint32_t x = s_wsum[i];
        base += i < w ? x : 0u;
        total += x;
    }
    uint32_t run = base + incl - sum;
#pragma unroll
    for (uint32_t k = 0; k < NL_PER; k++) {
        const uint32_t i = t * NL_PER + k;
        if (i < n) deg[i] = run;
        run += loc[k];
    }
    __syncthreads();   // s_wsum reuse, deg complete
    return total;
}

template <bool SINGLE, bool COUNT>
__global__ __launch_bounds__(NL_THREADS) void expand_narrow_kernel(DoArgs a)
{
    __shared__ uint32_t s_q[2][NL_CAP];
    __shared__ uint32_t s_beg[NL_CAP];
    __shared__ uint32_t s_off[NL_CAP + 1];
    __shared__ uint32_t s_cls[SINGLE ? 1 : NL_CAP];
    __shared__ uint32_t s_wsum[NL_THREADS / 64];
    __shared__ uint32_t s_n;
    const int dir = (int)blockIdx.x;
    if (dir >= a.ndirs) return;
    const DirArgs &D = a.d[dir];
    const int L0 = a.level;
    const uint32_t t = threadIdx.x;
    // the state the level-L0 launch would compute
    const DirState prev = load_dir_state(a.ctr, L0 == 0 ? 0 : (L0 - 1) & 1, dir);
    if (L0 > 0 && prev.mode != DIR_TD) return;   // finished, or a bottom-up level's slotted queue
    uint32_t n = D.cnt[CT(L0 % 3)];
    const uint32_t hubs = (uint32_t)(*D.pbigc >> 32);
    const DirState st0 = L0 > 0 ? dir_decide(prev, n, hubs, a.ctr[CT(C_NLIVE)], (uint32_t)L0, true) : prev;
    if (st0.mode != DIR_TD || hubs || n > NL_CAP) return;
    uint32_t *q = const_cast<uint32_t *>(D.qin);   // level L0's queue: read once, then rewritten
    for (uint32_t i = t; i < n; i += NL_THREADS) s_q[0][i] = q[i];
    __syncthreads();
    const uint32_t *__restrict__ R = D.R;
    const uint32_t *__restrict__ C = D.C;
    uint64_t visited = st0.visited;      // includes the current frontier
    uint32_t last_nf = prev.nf, levels = 0, cur = 0;
    uint32_t wc_v = 0, wc_e = 0, wc_c = 0;
    bool spilled = false;                // the frontier is in the global queue
    while (n) {
        // row offsets (and classes) of the frontier, every lane's loads in flight together
        uint32_t b[NL_PER], e[NL_PER], c[NL_PER];
#pragma unroll
        for (uint32_t k = 0; k < NL_PER; k++) {
            const uint32_t i = t * NL_PER + k;
            const uint32_t v = i < n ? s_q[cur][i] : 0u;
            b[k] = i < n ? R[v] : 0u;
            e[k] = i < n ? R[v + 1] : 0u;
            if (!SINGLE) c[k] = i < n ? a.cls[v] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < NL_PER; k++) {
            const uint32_t i = t * NL_PER + k;
            if (i < n) {
                s_beg[i] = b[k];
                s_off[i] = e[k] - b[k];
                if (!SINGLE) s_cls[i] = c[k];
            }
        }
        if (t == 0) s_n = 0;
        __syncthreads();
        const uint32_t T = nl_scan(s_off, n, s_wsum);
        if (levels > 0 && T > NL_ARCS) break;   // wide again: the device-wide levels take over
        if (t == 0) s_off[n] = T;
        __syncthreads();
        uint32_t *nq = s_q[cur ^ 1];
        for (uint32_t rb = 0; rb < T; rb += NL_THREADS * NL_K) {
            const uint32_t j0 = rb + t * NL_K;
            uint32_t w[NL_K], wcl[NL_K];
            uint32_t o = 0;
            if (j0 < T) {   // owner of arc j0: the last entry whose offset is <= j0
                uint32_t lo = 0, hi = n - 1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (s_off[mid] <= j0) lo = mid;
                    else hi = mid - 1;
                }
                o = lo;
            }
#pragma unroll
            for (uint32_t k = 0; k < NL_K; k++) {
                const uint32_t j = j0 + k;
                w[k] = 0;
                wcl[k] = 0;
                if (j < T) {
                    while (s_off[o + 1] <= j) o++;
                    w[k] = C[s_beg[o] + (j - s_off[o])];
                    if (!SINGLE) wcl[k] = s_cls[o];
                }
            }
            // every claim of the lane in flight at once, then the appends
            uint32_t old[NL_K];
#pragma unroll
            for (uint32_t k = 0; k < NL_K; k++) {
                const uint32_t j = j0 + k;
                old[k] = ~0u;
                if (j < T && (SINGLE || a.cls[w[k]] == wcl[k]))
                    old[k] = __hip_atomic_fetch_or(&D.vis[vword(w[k])], 1u << (w[k] & 31), __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            }
#pragma unroll
            for (uint32_t k = 0; k < NL_K; k++) {
                const uint32_t j = j0 + k;
                const bool claimed = !(old[k] & (1u << (w[k] & 31)));
                const uint64_t m = __ballot(claimed);
                if (m) {
                    const int leader = __ffsll((unsigned long long)m) - 1;
                    uint32_t base = 0;
                    if ((int)wlane() == leader) base = atomicAdd(&s_n, (uint32_t)__popcll(m));
                    base = __shfl(base, leader);
                    const uint32_t at = base + (uint32_t)__popcll(m & ((1ull << wlane()) - 1));
                    if (claimed) {
                        if (at < NL_CAP) nq[at] = w[k];
                        else q[at] = w[k];   // past the LDS frontier: straight to the global queue
                    }
                }
                if (COUNT) {
                    wc_e += j < T;
                    wc_c += claimed;
                }
            }
        }
        if (COUNT && t == 0) wc_v += n;
        __syncthreads();
        const uint32_t n2 = s_n;
        levels++;
        last_nf = n;
        visited += n2;
        n = n2;
        cur ^= 1;
        if (n > NL_SPILL) {   // the LDS part joins any claims already in the global queue
            for (uint32_t i = t; i < min(n, NL_CAP); i += NL_THREADS) q[i] = s_q[cur][i];
            spilled = true;
            break;
        }
        __syncthreads();   // s_n reuse
    }
    if (!spilled)
        for (uint32_t i = t; i < n; i += NL_THREADS) q[i] = s_q[cur][i];
    if (t == 0) {
        D.cnt[CT(L0 % 3)] = n;
        DirState s;
        s.mode = DIR_TD;
        s.end = 0;
        s.pad = 0;
        if (L0 == 0) {   // level 0 reads its state as it is
            s.nf = n;
            s.visited = visited;
            store_dir_state(a.ctr, 0, dir, s);
        } else {          // level L0 adds its input to `visited` itself
            s.nf = last_nf;
            s.visited = visited - n;
            store_dir_state(a.ctr, (L0 - 1) & 1, dir, s);
        }
        atomicAdd(&a.ctr[CT(C_NLEVELS) + dir], levels);
    }
    if (COUNT) {
        WorkCount wc;
        wc.n[W_TDV] = wc_v;
        wc.n[W_TDE] = wc_e;
        wc.n[W_CLAIM] = wc_c;
        add_work(a.ctr, wc);
    }
}

__global__ void do_begin_kernel(uint32_t *ctr, int ndirs)
{
    const int d = (int)threadIdx.x;
    if (d >= 2) return;
    DirState st;
    st.mode = d < ndirs ? DIR_TD : DIR_OFF;
    st.end = 0;
    st.nf = ctr[CT(C_NPIV)];
    st.pad = 0;
    st.visited = ctr[CT(C_NPIV)];
    store_dir_state(ctr, 0, d, st);
}

void k_narrow_levels(hipStream_t s, const DoArgs &a, bool single_class)
{
    if (single_class && a.count)
        expand_narrow_kernel<true, true><<<a.ndirs, NL_THREADS, 0, s>>>(a);
    else if (single_class)
        expand_narrow_kernel<true, false><<<a.ndirs, NL_THREADS, 0, s>>>(a);
    else
        expand_narrow_kernel<false, false><<<a.ndirs, NL_THREADS, 0, s>>>(a);
}

uint32_t narrow_cap() { return NL_ENTER; }

uint32_t do_blocks(uint32_t V, int ndirs)
{
    const uint64_t b = ((uint64_t)V + 256) / 256;
    return dir_blocks((uint32_t)(b > 2048 ? 2048 : b), ndirs);
}

size_t queue_words(uint32_t V)
{
    size_t w = (size_t)V + 1;
    for (int nd = 1; nd <= 2; nd++) w = std::max(w, (size_t)kSpread * slot_cap(V, do_blocks(V, nd)));
    return w;
}

void k_do_begin(hipStream_t s, uint32_t *ctr, int ndirs) { do_begin_kernel<<<1, 64, 0, s>>>(ctr, ndirs); }

void k_do_level(hipStream_t s, const DoArgs &a, bool single_class, uint32_t bpd)
{
    bpd = dir_blocks(bpd, a.ndirs);
    const uint32_t grid = bpd * (uint32_t)a.ndirs;
    if (single_class && a.count)
        expand_do_kernel<true, true><<<grid, 256, 0, s>>>(a, bpd);
    else if (single_class)
        expand_do_kernel<true, false><<<grid, 256, 0, s>>>(a, bpd);
    else
        expand_do_kernel<false, false><<<grid, 256, 0, s>>>(a, bpd);
}

template <int W>
static void launch_vw(hipStream_t s, const LevelArgs &a, bool single, uint32_t bpd)
{
    bpd = dir_blocks(bpd, a.ndirs);
    const uint32_t grid = bpd * (uint32_t)a.ndirs;
    if (single)
        expand_vw_kernel<W, true><<<grid, 256, 0, s>>>(a, bpd);
    else
        expand_vw_kernel<W, false><<<grid, 256, 0, s>>>(a, bpd);
}

void k_expand_level(hipStream_t s, const LevelArgs &a, int expand, int warp, bool single_class,
                    uint32_t blocks_per_dir)
{
    if (expand == SCC_EXPAND_LB) {
        launch_lb(s, a, single_class, blocks_per_dir);
        return;
    }
    int W = expand == SCC_EXPAND_VWARP ? warp : 1;
    switch (W) {
    case 1: launch_vw<1>(s, a, single_class, blocks_per_dir); break;
    case 2: launch_vw<2>(s, a, single_class, blocks_per_dir); break;
    case 4: launch_vw<4>(s, a, single_class, blocks_per_dir); break;
    case 8: launch_vw<8>(s, a, single_class, blocks_per_dir); break;
    case 16: launch_vw<16>(s, a, single_class, blocks_per_dir); break;
    case 32: launch_vw<32>(s, a, single_class, blocks_per_dir); break;
    default: launch_vw<64>(s, a, single_class, blocks_per_dir); break;
    }
}

} // namespace sccamd
