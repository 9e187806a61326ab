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
    if (bid == 0) seg_clear(&D.cnt[CT((L + 2) % 3)], &D.cslot[CT(((L + 2) % 3) * kSpread)]);
    if (n == 0) return;
    const uint32_t g = (bid * blockDim.x + threadIdx.x) / W;
    const uint32_t ng = bpd * blockDim.x / W;
    const uint32_t sub = threadIdx.x % W;
    const uint32_t *__restrict__ R = D.R;
    const uint32_t *__restrict__ C = D.C;
    for (uint32_t i = g; i < n; i += ng) {
        const uint32_t v = D.qin[i];
        const uint32_t c = SINGLE ? 0u : a.cls[v];
        const uint32_t e = R[v + 1];
        for (uint32_t j = R[v] + sub; j < e; j += W) {
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
template <bool SINGLE>
__device__ __forceinline__ void lb_level(const DirArgs &D, const uint32_t *cls, int L, uint32_t bid,
                                         uint32_t bpd, uint32_t *lds, uint32_t n, const uint32_t *pre,
                                         uint32_t qcap)
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
            }
            lq_push(q, claimed, w);
            lq_flush(q, cnt_out, D.qout, false);
        }
        __syncthreads();   // LDS reuse by the next tile
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
    lb_level<SINGLE>(D, a.cls, a.level, bid, bpd, s_lds, D.cnt[CT(a.level % 3)], nullptr, 0u);
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

// One 64-bit load gives p's visited-or-finished and finished bits.
template <bool SINGLE>
__device__ __forceinline__ bool visited_parent(const uint32_t *vis, const uint32_t *cls, uint32_t p, uint32_t c)
{
    const uint64_t w = reinterpret_cast<const uint64_t *>(vis)[p >> 5];
    const uint32_t pb = 1u << (p & 31);
    return ((uint32_t)w & ~(uint32_t)(w >> 32) & pb) && (SINGLE || cls[p] == c);
}

// LDS words of a bottom-up level: each wave's queue, long list and the
// long-list hits of its chunks (two words per chunk).
constexpr uint32_t BU_LDS = 4 * BU_WQ + 4 * BU_CHUNKS * 64 + 4 * BU_CHUNKS * 2;

// One bottom-up level of direction D (parents in PR/PC) for workgroup `bid`
// of `bpd`; `lds` holds BU_LDS words.  Called by the whole workgroup.
//
// The vertices a pass visits go to the region of slot bid % kSpread of the
// output queue (slot_cap(V, bpd) = qcap entries: the slot's workgroups own
// that many vertices at most) and are counted in that slot of the count, so
// the flushes of the pass's workgroups are spread over kSpread addresses.
template <bool SINGLE>
__device__ __forceinline__ void bu_level(const DirArgs &D, const uint32_t *__restrict__ PR,
                                         const uint32_t *__restrict__ PC, const uint32_t *__restrict__ PF,
                                         const uint32_t *cls, uint32_t V, int L, uint32_t bid, uint32_t bpd,
                                         uint32_t *lds, uint32_t qcap, uint32_t *anom)
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
    uint32_t *wh = lds + 4 * BU_WQ + 4 * BU_CHUNKS * 64 + wv * BU_CHUNKS * 2;
    uint32_t nq = 0;
    const uint32_t sub = lane % BU_GROUP, grp = lane / BU_GROUP;
    const uint32_t gshift = lane & ~(BU_GROUP - 1);
    const uint32_t stride = bpd * 256;   // one chunk of every wave of the grid
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
        // parent misses loads its row offsets to see whether more follow
        uint32_t pf[BU_CHUNKS];
#pragma unroll
        for (uint32_t h = 0; h < BU_CHUNKS; h++) {
            c[h] = (SINGLE || !cand[h]) ? 0u : cls[u[h]];
            pf[h] = cand[h] ? PF[u[h]] : 0u;
        }
#pragma unroll
        for (uint32_t h = 0; h < BU_CHUNKS; h++) {
            found[h] = pf[h] && visited_parent<SINGLE>(D.vis, cls, pf[h], c[h]);
            hits[h] = __ballot(found[h]);
            const bool more = pf[h] && !found[h];
            b[h] = more ? PR[u[h]] : 0u;
            e[h] = more ? PR[u[h] + 1] : 0u;
            need_long[h] = more && e[h] - b[h] > BU_SHORT;
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
        // long parent lists: 64 / BU_GROUP vertices per round, each group
        // stops at its first hit
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
                    hit = ((__ballot(h) >> gshift) & ((1ull << BU_GROUP) - 1)) != 0;
                }
            }
            const bool claimed = hit && sub == 0;
            if (claimed) {
                const uint32_t hx = (x - base) / stride, off = x - base - hx * stride;
                atomicOr(&wh[hx * 2 + (off >> 5)], 1u << (off & 31));   // LDS
            }
            wave_push(wq, nq, claimed, x);
        }
        wave_sync();   // long-list hits are in wh; the next iteration rewrites wl
#pragma unroll
        for (uint32_t h = 0; h < BU_CHUNKS; h++) {
            // lanes 0 and 32 own the chunk's two visited words
            const uint32_t m = (lane == 0 ? (uint32_t)hits[h] : (uint32_t)(hits[h] >> 32)) | wh[h * 2 + (lane >> 5)];
            if ((lane & 31) == 0 && m) D.vis[vword(u[h])] = word[h] | m;
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
template <bool SINGLE>
__global__ __launch_bounds__(256) void expand_do_kernel(DoArgs a, uint32_t bpd)
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
    if (st.mode == DIR_TD)
        lb_level<SINGLE>(D, a.cls, L, bid, bpd, s_lds, n, slotted ? s_pre : nullptr, a.qcap);
    else if (st.mode == DIR_BU)
        bu_level<SINGLE>(D, a.PR[dir], a.PC[dir], a.PF[dir], a.cls, a.V, L, bid, bpd, s_lds, a.qcap,
                         a.ctr + CT(C_ANOM));
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
    if (single_class)
        expand_do_kernel<true><<<grid, 256, 0, s>>>(a, bpd);
    else
        expand_do_kernel<false><<<grid, 256, 0, s>>>(a, bpd);
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
