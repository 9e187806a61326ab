// kernels_sweep.hip -- vertex-sweep kernels of the SCC engine for gfx950.
//
// Every kernel walks vertices 1..V in wave-uniform strides: a wave owns 64
// consecutive vertices per step, so ballots and cross-lane reductions always
// see a full wave.  Class and state updates that other lanes may race on use
// atomics (the reference's byte read-modify-writes, scc.h:28-42, are not).
#include <algorithm>

#include "lds_queue.h"
#include "scc_dev.h"

namespace sccamd {

#define WAVE_LOOP(v, valid, V)                                                              \
    for (uint32_t _base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); _base < (V);     \
         _base += gridDim.x * blockDim.x)                                                   \
        for (uint32_t v = _base + (threadIdx.x & 63u) + 1, _once = 1; _once; _once = 0)     \
            for (bool valid = v <= (V); _once; _once = 0)

__device__ __forceinline__ uint32_t lane() { return threadIdx.x & 63u; }

// pivot key of v from its out-degree and in-degree
__device__ __forceinline__ uint64_t pivot_key_deg(uint32_t outdeg, uint32_t indeg, uint32_t v)
{
    uint64_t p = (uint64_t)outdeg * (uint64_t)indeg;
    p = p > 0xFFFFFFFFull ? 0xFFFFFFFFull : p;
    return (p << 32) | (uint64_t)(0xFFFFFFFFu - v);
}

__device__ __forceinline__ uint64_t pivot_key(const uint32_t *__restrict__ Fr,
                                              const uint32_t *__restrict__ Br, uint32_t v)
{
    return pivot_key_deg(Fr[v + 1] - Fr[v], Br[v + 1] - Br[v], v);
}

__device__ __forceinline__ uint32_t key_vertex(uint64_t k) { return 0xFFFFFFFFu - (uint32_t)k; }

// visited (or finished) bit of v in fvis / bvis
__device__ __forceinline__ bool bit_of(const uint32_t *vis, uint32_t v)
{
    return (vis[vword(v)] >> (v & 31)) & 1u;
}

// Sum over the whole workgroup, added to *p by one atomic.  Every thread of
// the block must call it (they all do: it ends the kernels that use it).
__device__ __forceinline__ uint64_t block_sum(uint64_t x)
{
    __shared__ uint64_t s_part[16];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if (lane() == 0) s_part[threadIdx.x >> 6] = x;
    __syncthreads();
    uint64_t t = 0;
    if (threadIdx.x == 0)
        for (uint32_t w = 0; w < (blockDim.x >> 6); w++) t += s_part[w];
    __syncthreads();
    return t;   // valid in thread 0
}

__device__ __forceinline__ void wave_add64(uint32_t *p32, uint64_t x)
{
    const uint64_t t = block_sum(x);
    if (threadIdx.x == 0 && t) atomicAdd(reinterpret_cast<unsigned long long *>(p32), (unsigned long long)t);
}

// Workgroup sum of x added to slot blockIdx.x % kSpread of spread counter
// c: the workgroups of a sweep end together, and one address taking an
// atomic from each of them would serialise the sweep's tail.
__device__ __forceinline__ void spread_add32(uint32_t *ctr, int c, uint32_t x)
{
    const uint64_t t = block_sum(x);
    if (threadIdx.x == 0 && t) atomicAdd(&ctr[CSLOT(c, blockIdx.x % kSpread)], (uint32_t)t);
}

uint32_t sweep_blocks(uint32_t V)
{
    uint64_t b = ((uint64_t)V + SWEEP_BLOCK) / SWEEP_BLOCK;
    // 2,048 workgroups (every wave slot of the chip) above 1M vertices: the
    // sweeps' dependent loads need them all (a plain streaming sweep runs at
    // 3.3 TB/s with 1,024 and 5.2 TB/s with 2,048).  The workgroups' closing
    // atomics go to spread counters, so more workgroups cost no serialised
    // tail (measured: R-MAT scale 22 0.409 -> 0.389 ms per decomposition)
    const uint64_t cap = V > (1u << 20) ? 2048 : 1024;
    return (uint32_t)(b < 1 ? 1 : (b > cap ? cap : b));
}

// ------------------------------------------------------------ init ---

__global__ void fill_u32_kernel(uint32_t *p, uint32_t val, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = val;
}

void k_fill_u32(hipStream_t s, uint32_t *p, uint32_t val, uint64_t n)
{
    uint64_t b = (n + SWEEP_BLOCK - 1) / SWEEP_BLOCK;
    fill_u32_kernel<<<(uint32_t)(b < 1 ? 1 : (b > 4096 ? 4096 : b)), SWEEP_BLOCK, 0, s>>>(p, val, n);
}

__global__ void first_arcs_kernel(const uint32_t *__restrict__ Fr, const uint32_t *__restrict__ Fc,
                                  const uint32_t *__restrict__ Br, const uint32_t *__restrict__ Bc, uint32_t *F1,
                                  uint32_t *B1, uint32_t *DP, uint32_t V)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= V; v += gridDim.x * blockDim.x) {
        const uint32_t fb = Fr[v], fe = Fr[v + 1], bb = Br[v], be = Br[v + 1];
        F1[v] = fb < fe ? Fc[fb] : 0u;
        B1[v] = bb < be ? Bc[bb] : 0u;
        const uint64_t p = (uint64_t)(fe - fb) * (uint64_t)(be - bb);
        DP[v] = p > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)p;
    }
}

void k_first_arcs(hipStream_t s, const DevGraph &g)
{
    first_arcs_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(g.Fr, g.Fc, g.Br, g.Bc, g.F1, g.B1, g.DP, g.V);
}

// All vertices start live in class 0 (range 0, scc_WCC.cu:630-631) unless
// `dead` marks them finished; lab[v] = v.  With vertices == 0 only the
// counters and the dummy vertex 0 are written: the first trim-1 pass of a
// fresh state writes every other vertex's words itself.
__global__ void init_state_kernel(uint32_t *cls, uint32_t *lab, uint64_t *piv, const uint8_t *dead,
                                  uint32_t *ctr, uint32_t V, int vertices)
{
    uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid == 0) {
        cls[0] = DEAD;
        lab[0] = 0;
        piv[0] = 0;
    }
    for (uint32_t i = tid; i < (uint32_t)kCtrLines; i += gridDim.x * blockDim.x) {
        ctr[CT(i)] = 0;   // both words of every counter line (64-bit counters too)
        ctr[CT(i) + 1] = 0;
    }
    for (uint32_t v = tid + 1; vertices && v <= V; v += gridDim.x * blockDim.x) {
        cls[v] = (dead && dead[v]) ? DEAD : 0u;
        lab[v] = v | LAB_FINAL;
    }
}

void k_init_state(hipStream_t s, const DevGraph &g, DevState &st, const uint8_t *dead_dev, bool vertices)
{
    init_state_kernel<<<vertices ? sweep_blocks(g.V) : 1u, SWEEP_BLOCK, 0, s>>>(st.cls, st.lab, st.piv, dead_dev,
                                                                               st.ctr, g.V, vertices ? 1 : 0);
}

// ---------------------------------------------------------- trim-1 ---
// scc_kernels.cu:187-233: a live vertex with no live same-class
// in-neighbour, or no live same-class out-neighbour, is a singleton SCC.
//
// The first pass sweeps every vertex; each later pass re-checks only the
// live neighbours of the vertices the previous pass removed (a vertex can
// only become removable when a neighbour goes), so the fixpoint costs one
// sweep plus work proportional to the removals (while a pass removes a
// large share of the vertices, the next one sweeps again: walking the arcs
// of millions of fresh removals costs more than a sweep).  In a list pass a
// candidate is tested once (the first lane to stamp it with the pass's
// epoch) and removal is a CAS on the class word, so each removal is listed
// once; in a sweep each vertex has one lane.  Pass k reads list count
// C_TLIST + k%3, appends to C_TLIST + (k+1)%3 and clears C_TLIST + (k+2)%3.

// Is w live and in class c?  With BITS every live vertex is in one class
// (the first trim of a decomposition) and the answer is one bit of the
// live bitmap `live` (1 bit per vertex, 2 MB at 16M vertices: it stays in
// L2, where the class words -- 4 bytes per vertex -- do not); otherwise it
// is a class compare.
template <bool BITS>
__device__ __forceinline__ bool same_live(const uint32_t *cls, const uint32_t *live, uint32_t w, uint32_t c)
{
    if (BITS) return (live[w >> 5] >> (w & 31)) & 1u;
    return cls[w] == c;
}

// Whether any of the arcs col[i..e) leads to a live same-class vertex.  A
// walk step loads TRIM_SCAN columns and then probes them all (indexes past
// the end repeat the last arc), so a vertex whose neighbours were mostly
// removed does not hold its wave for two dependent loads per arc.
constexpr uint32_t TRIM_SCAN = 4;

template <bool BITS>
__device__ __forceinline__ bool any_live(const uint32_t *__restrict__ col, uint32_t i, uint32_t e,
                                         const uint32_t *cls, const uint32_t *live, uint32_t c)
{
    bool hit = false;
    for (; !hit && i < e; i += TRIM_SCAN) {
        uint32_t w[TRIM_SCAN];
#pragma unroll
        for (uint32_t k = 0; k < TRIM_SCAN; k++) w[k] = col[min(i + k, e - 1)];
#pragma unroll
        for (uint32_t k = 0; k < TRIM_SCAN; k++) hit |= same_live<BITS>(cls, live, w[k], c);
    }
    return hit;
}

// v (live, class c) keeps a live same-class in-neighbour and out-neighbour.
// The first arc each way is probed first: pin = B1[v] and pout = F1[v],
// loaded with the row offsets (coalesced, one dependent step to the
// neighbours' class words or live bits instead of a random list line each),
// and only a vertex with no hit there loads its row offsets and walks the
// rest of a list (a first neighbour of 0 means an empty list: vertex ids
// start at 1).  Probing more arcs up front adds random lines and is slower.
template <bool BITS>
__device__ __forceinline__ bool trim1_keep_at(const uint32_t *__restrict__ Fr, const uint32_t *__restrict__ Fc,
                                              const uint32_t *__restrict__ Br, const uint32_t *__restrict__ Bc,
                                              const uint32_t *cls, const uint32_t *live, uint32_t c, uint32_t v,
                                              uint32_t pin, uint32_t pout)
{
    // vertex 0 (the first neighbour of an empty list) is dead in the class
    // words and the live bitmap, so both probes go out unconditionally
    bool kin = same_live<BITS>(cls, live, pin, c);
    bool kout = same_live<BITS>(cls, live, pout, c);
    if (!kin && pin) kin = any_live<BITS>(Bc, Br[v] + 1, Br[v + 1], cls, live, c);
    if (kin && !kout && pout) kout = any_live<BITS>(Fc, Fr[v] + 1, Fr[v + 1], cls, live, c);
    return kin && kout;
}

// One pass of the fixpoint.  Pass 0, and any pass whose input list holds
// more than V / sweep_share vertices, sweeps every vertex; the others test
// the live neighbours of the listed removals, TRIM_W lanes per removed
// vertex and one arc per lane per round (the workgroup runs as many rounds
// as its longest list needs).  The pass reads list count C_TLIST + pass%3,
// appends to C_TLIST + (pass+1)%3 and clears C_TLIST + (pass+2)%3, so the
// host can queue several passes: once a pass removes nothing the following
// ones find an empty list and only clear counters.  Removals go to
// wave-private LDS queues (a sweep can remove millions of vertices).
//
// A list count is a spread counter: a sweep reserves list space from slot
// blockIdx.x % kSpread and writes into that slot's region of the list (the
// slot's workgroups own a fixed share of the 64-vertex chunks, which bounds
// the region: trim_slot_cap), a list pass reserves from the counter's own
// word and writes from index 0.  A pass only ever writes one way, and the
// next pass reads the segments (own word first, then the slots) through
// their prefix sums.
//
// A removed vertex keeps its label: a live vertex's label is still the
// v | LAB_FINAL the state was initialised with, which is its trim-1 label.
// A sweep gives each wave 64-aligned chunks, so the wave owns its chunk's
// class words and (BITS) its two live-bitmap words: removal is plain
// stores.  A list pass may test one vertex from several lanes: the
// first to stamp it with the pass's epoch tests it, removal is a CAS on its
// class word and (BITS) an atomic clear of its live bit.
constexpr uint32_t TRIM_WQ = 1024;
constexpr uint32_t TRIM_W = 8;
constexpr uint32_t TRIM_PER_BLOCK = 256 / TRIM_W;

// Largest share of a sweep's removals one slot can take: its workgroups
// (every kSpread-th of sweep_blocks(V)) hold 4 waves each, and a wave owns
// at most ceil((V + 1) / (256 * grid)) chunks of 64 vertices.
uint32_t slot_cap(uint32_t V, uint32_t grid_)
{
    const uint64_t grid = grid_ < 1 ? 1 : grid_;
    const uint64_t per_wave = ((uint64_t)V + 1 + 256 * grid - 1) / (256 * grid);
    return (uint32_t)(((grid + kSpread - 1) / kSpread) * 256 * per_wave);
}

uint32_t trim_slot_cap(uint32_t V) { return slot_cap(V, sweep_blocks(V)); }

size_t trim_list_words(uint32_t V)
{
    const size_t slotted = (size_t)kSpread * trim_slot_cap(V);
    return std::max<size_t>(slotted, (size_t)V + 1);
}

template <bool BITS>
__global__ __launch_bounds__(256) void trim1_kernel(const uint32_t *__restrict__ Fr,
                                                    const uint32_t *__restrict__ Fc,
                                                    const uint32_t *__restrict__ Br,
                                                    const uint32_t *__restrict__ Bc,
                                                    const uint32_t *__restrict__ F1,
                                                    const uint32_t *__restrict__ B1, uint32_t *cls,
                                                    uint32_t *lab, uint32_t *live, uint32_t V, uint32_t *ctr,
                                                    const uint32_t *in, uint32_t *out, uint32_t *stamp,
                                                    uint32_t epoch, uint32_t pass, uint32_t sweep_share,
                                                    uint32_t cap, int lazy)
{
    __shared__ uint32_t s_q[4][TRIM_WQ];
    __shared__ uint32_t s_rounds;
    __shared__ uint32_t s_pre[kSpread + 2];
    const uint32_t wv = threadIdx.x >> 6;
    uint32_t *wq = s_q[wv];
    uint32_t nq = 0, removed = 0;
    const int c_in = C_TLIST + pass % 3, c_out = C_TLIST + (pass + 1) % 3;
    if (pass) seg_prefix(&ctr[CT(c_in)], &ctr[CSLOT(c_in, 0)], cap, s_pre);
    const uint32_t n = pass ? s_pre[kSpread + 1] : 0u;
    if (blockIdx.x == 0) seg_clear(&ctr[CT(C_TLIST + (pass + 2) % 3)], &ctr[CSLOT(C_TLIST + (pass + 2) % 3, 0)]);
    const bool sweep = pass == 0 || n > V / sweep_share;
    const uint32_t slot = blockIdx.x % kSpread;
    uint32_t *cnt_out = sweep ? &ctr[CSLOT(c_out, slot)] : &ctr[CT(c_out)];
    uint32_t *qout = sweep ? out + (size_t)slot * cap : out;
    const uint32_t qcap = sweep ? cap : ~0u;   // a list pass removes each vertex once: at most V entries
    uint32_t *anom = &ctr[CT(C_ANOM)];
    if (sweep) {
        for (uint32_t base = (blockIdx.x * 4 + wv) * 64; base <= V; base += gridDim.x * 256) {
            if (nq + 64 > TRIM_WQ) wave_flush(wq, nq, cnt_out, qout, qcap, anom);
            const uint32_t v = base + lane();   // vertex 0, the dummy, is dead
            // the first neighbours are loaded beside the live word or class
            // (coalesced; a dead vertex's entries are not used, and past V
            // the lane reads vertex 0's)
            const uint32_t vv = v <= V ? v : 0u;
            const uint32_t pin = B1[vv], pout = F1[vv];
            uint32_t word = 0, c = DEAD;
            if (v <= V) {
                if (BITS) {
                    word = live[v >> 5];
                    c = (word >> (v & 31)) & 1u ? 0u : DEAD;
                } else {
                    c = cls[v];
                }
            }
            bool gone = false;
            if (c != DEAD) {
                if (!trim1_keep_at<BITS>(Fr, Fc, Br, Bc, cls, live, c, v, pin, pout)) {
                    if (!lazy) cls[v] = DEAD;
                    gone = true;
                }
            }
            if (BITS) {
                const uint64_t m = __ballot(gone);
                const uint32_t mw = lane() < 32 ? (uint32_t)m : (uint32_t)(m >> 32);
                if ((lane() & 31) == 0 && mw) live[v >> 5] = word & ~mw;
            }
            removed += gone;
            wave_push(wq, nq, gone, v);
        }
    } else {
        const uint32_t sub = threadIdx.x % TRIM_W;
        for (uint32_t base = blockIdx.x * TRIM_PER_BLOCK; base < n; base += gridDim.x * TRIM_PER_BLOCK) {
            const uint32_t i = base + threadIdx.x / TRIM_W;
            uint32_t fb = 0, nf = 0, bb = 0, total = 0;
            if (i < n) {
                const uint32_t v = in[seg_pos(s_pre, i, cap)];
                fb = Fr[v];
                nf = Fr[v + 1] - fb;
                bb = Br[v];
                total = nf + (Br[v + 1] - bb);
            }
            if (threadIdx.x == 0) s_rounds = 0;
            __syncthreads();
            if (sub == 0 && total) atomicMax(&s_rounds, (total + TRIM_W - 1) / TRIM_W);
            __syncthreads();
            const uint32_t rounds = s_rounds;
            for (uint32_t r = 0; r < rounds; r++) {
                if (nq + 64 > TRIM_WQ) wave_flush(wq, nq, cnt_out, qout, qcap, anom);
                const uint32_t j = r * TRIM_W + sub;
                bool gone = false;
                uint32_t w = 0;
                if (j < total) {
                    w = j < nf ? Fc[fb + j] : Bc[bb + (j - nf)];
                    const uint32_t c = BITS ? ((live[w >> 5] >> (w & 31)) & 1u ? 0u : DEAD) : cls[w];
                    if (c != DEAD && stamp[w] != epoch && atomicExch(&stamp[w], epoch) != epoch &&
                        !trim1_keep_at<BITS>(Fr, Fc, Br, Bc, cls, live, c, w, B1[w], F1[w]) &&
                        (BITS ? (atomicAnd(&live[w >> 5], ~(1u << (w & 31))) >> (w & 31)) & 1u
                              : atomicCAS(&cls[w], c, DEAD) == c)) {
                        if (BITS && !lazy) cls[w] = DEAD;
                        gone = true;
                    }
                }
                removed += gone;
                wave_push(wq, nq, gone, w);
            }
            __syncthreads();   // s_rounds reuse
        }
    }
    wave_flush_block(wq, nq, cnt_out, qout, qcap, anom);
    spread_add32(ctr, C_TRIM1, removed);
}

// Pass 0 of a single-class fixpoint from a fresh state (no dead vertex,
// every vertex in class 0 with label v | LAB_FINAL), fused with writing that
// state and building the live bitmap: a vertex without an in- or out-arc is
// removed (class DEAD, listed in `out`), every other vertex gets class 0 and
// its live bit, and every vertex gets label v | LAB_FINAL, so the state
// initialisation before it may skip the per-vertex words.  Without
// write_labels (lazy) the labels and class words are left unwritten and the
// whole single-class trim works from the live bitmap: a decomposition that
// the first round finishes computes every label from the search's bitmaps
// (label_from_bits_kernel), and otherwise label_fixup_kernel writes the
// labels and class words the eager passes would have.
// The pass also builds the per-vertex arrays the later phases read, for
// the vertices it keeps (nothing reads them for a removed vertex): the
// degree product DP (the first pivot's election key), and the first out-
// and in-neighbour F1/B1 (the trim-1 and bottom-up first-arc probes).  So
// it streams the row offsets (coalesced; 0 degree product: no out-arc or no
// in-arc) and gathers the first column entry of each kept vertex's two
// lists, with TRIM_DEG_CHUNKS 64-vertex chunks' loads in flight per wave.
constexpr uint32_t TRIM_DEG_CHUNKS = 4;

__global__ __launch_bounds__(256) void trim1_degree_kernel(const uint32_t *__restrict__ Fr,
                                                           const uint32_t *__restrict__ Fc,
                                                           const uint32_t *__restrict__ Br,
                                                           const uint32_t *__restrict__ Bc, uint32_t *F1,
                                                           uint32_t *B1, uint32_t *DP, uint32_t *cls,
                                                           uint32_t *lab, uint32_t *live, uint32_t V,
                                                           uint32_t *ctr, uint32_t *out, uint32_t cap,
                                                           int write_labels)
{
    __shared__ uint32_t s_q[4][TRIM_WQ];
    const uint32_t wv = threadIdx.x >> 6;
    uint32_t *wq = s_q[wv];
    uint32_t nq = 0, removed = 0;
    const uint32_t slot = blockIdx.x % kSpread;
    uint32_t *cnt_out = &ctr[CSLOT(C_TLIST + 1, slot)];
    uint32_t *qout = out + (size_t)slot * cap;
    uint32_t *anom = &ctr[CT(C_ANOM)];
    if (blockIdx.x == 0) seg_clear(&ctr[CT(C_TLIST + 2)], &ctr[CSLOT(C_TLIST + 2, 0)]);
    const uint32_t stride = gridDim.x * 256;
    for (uint32_t base = (blockIdx.x * 4 + wv) * 64; base <= V; base += TRIM_DEG_CHUNKS * stride) {
        if (nq + TRIM_DEG_CHUNKS * 64 > TRIM_WQ) wave_flush(wq, nq, cnt_out, qout, cap, anom);
        uint32_t v[TRIM_DEG_CHUNKS], fb[TRIM_DEG_CHUNKS], fe[TRIM_DEG_CHUNKS], bb[TRIM_DEG_CHUNKS],
            be[TRIM_DEG_CHUNKS], dp[TRIM_DEG_CHUNKS];
#pragma unroll
        for (uint32_t h = 0; h < TRIM_DEG_CHUNKS; h++) {
            v[h] = base + h * stride + lane();
            const bool in = v[h] <= V;
            fb[h] = in ? Fr[v[h]] : 0u;
            fe[h] = in ? Fr[v[h] + 1] : 0u;
            bb[h] = in ? Br[v[h]] : 0u;
            be[h] = in ? Br[v[h] + 1] : 0u;
        }
        // every chunk's first-neighbour gathers are issued before any is
        // stored (a gather and its store per branch would wait for each)
        uint32_t f1[TRIM_DEG_CHUNKS], b1[TRIM_DEG_CHUNKS];
#pragma unroll
        for (uint32_t h = 0; h < TRIM_DEG_CHUNKS; h++) {
            // 0: no out-arc or no in-arc (vertex 0, the dummy, has neither)
            const uint64_t p = (uint64_t)(fe[h] - fb[h]) * (uint64_t)(be[h] - bb[h]);
            dp[h] = p > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)p;
            f1[h] = 0;
            b1[h] = 0;
            if (v[h] >= 1 && dp[h]) {
                f1[h] = Fc[fb[h]];
                b1[h] = Bc[bb[h]];
            }
        }
#pragma unroll
        for (uint32_t h = 0; h < TRIM_DEG_CHUNKS; h++) {
            if (v[h] >= 1 && dp[h]) {
                F1[v[h]] = f1[h];
                B1[v[h]] = b1[h];
                DP[v[h]] = dp[h];
            }
        }
#pragma unroll
        for (uint32_t h = 0; h < TRIM_DEG_CHUNKS; h++) {
            const bool alive = v[h] >= 1 && v[h] <= V;   // vertex 0, the dummy, is dead
            const bool gone = alive && dp[h] == 0;
            if (alive && write_labels) {
                cls[v[h]] = gone ? DEAD : 0u;
                lab[v[h]] = v[h] | LAB_FINAL;
            }
            const uint64_t m = __ballot(alive && !gone);
            if ((lane() & 31) == 0 && v[h] <= V) live[v[h] >> 5] = lane() < 32 ? (uint32_t)m : (uint32_t)(m >> 32);
            removed += gone;
            wave_push(wq, nq, gone, v[h]);
        }
    }
    wave_flush_block(wq, nq, cnt_out, qout, cap, anom);
    spread_add32(ctr, C_TRIM1, removed);
}

// The live bitmap of the single-class trim: bit v set for every vertex
// whose class word is not DEAD (vertex 0 included, as dead).
__global__ void live_bits_kernel(const uint32_t *__restrict__ cls, uint32_t *live, uint32_t V)
{
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t base = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64; base <= V; base += nwaves * 64) {
        const uint32_t v = base + lane();
        const uint64_t m = __ballot(v <= V && cls[v] != DEAD);
        if ((lane() & 31) == 0) live[v >> 5] = lane() < 32 ? (uint32_t)m : (uint32_t)(m >> 32);
    }
}

void k_trim1(hipStream_t s, const DevGraph &g, DevState &st, const uint32_t *in, uint32_t *out,
             uint32_t epoch, uint32_t pass, uint32_t sweep_share, uint32_t *live, bool degree_first, bool labels)
{
    const uint32_t cap = trim_slot_cap(g.V);
    if (live && pass == 0 && degree_first) {
        trim1_degree_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(g.Fr, g.Fc, g.Br, g.Bc, g.F1, g.B1, g.DP,
                                                                      st.cls, st.lab, live, g.V, st.ctr, out,
                                                                      cap, labels ? 1 : 0);
    } else if (live) {
        if (pass == 0)
            live_bits_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(st.cls, live, g.V);
        trim1_kernel<true><<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(g.Fr, g.Fc, g.Br, g.Bc, g.F1, g.B1, st.cls,
                                                                     st.lab,
                                                                     live, g.V, st.ctr, in, out, st.stamp,
                                                                     epoch, pass, sweep_share, cap, labels ? 0 : 1);
    } else {
        trim1_kernel<false><<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(g.Fr, g.Fc, g.Br, g.Bc, g.F1, g.B1,
                                                                      st.cls, st.lab,
                                                                      nullptr, g.V, st.ctr, in, out,
                                                                      st.stamp, epoch, pass, sweep_share, cap, 0);
    }
}

// ---------------------------------------------------------- trim-2 ---
// scc_kernels.cu:236-357.  Live same-class degree capped at 2; *k gets
// the last live neighbour seen.
// Two arcs per step (both columns, then both class tests): the answer
// usually rests on the first two arcs, so a step of two halves the chain
// of dependent loads without reading arcs that are not needed; an index
// past the end repeats the last arc and is not counted.  With BITS (every
// live vertex in one class) the test is the live bitmap, which stays in L2,
// instead of a random line of the class words.
template <bool BITS>
__device__ __forceinline__ int live_deg2(const uint32_t *__restrict__ R, const uint32_t *__restrict__ C,
                                         const uint32_t *cls, const uint32_t *live, uint32_t v, uint32_t c,
                                         uint32_t *k)
{
    int d = 0;
    for (uint32_t i = R[v], e = R[v + 1]; i < e; i += 2) {
        const uint32_t w0 = C[i], w1 = C[min(i + 1, e - 1)];
        const bool s0 = same_live<BITS>(cls, live, w0, c), s1 = same_live<BITS>(cls, live, w1, c);
        if (s0) {
            if (++d == 2) return 2;
            *k = w0;
        }
        if (i + 1 < e && s1) {
            if (++d == 2) return 2;
            *k = w1;
        }
    }
    return d;
}

__device__ __forceinline__ bool has_arc(const uint32_t *__restrict__ R,
                                        const uint32_t *__restrict__ C, uint32_t v, uint32_t k)
{
    for (uint32_t i = R[v], e = R[v + 1]; i < e; i++)
        if (C[i] == k) return true;
    return false;
}

// A pair {v,k} whose only live same-class in-arcs (or out-arcs) are the
// ones between them is an SCC.  The claim is a CAS on the smaller id, so the
// pair is recorded once and no flag update is lost (defect D5).  With BITS
// every live vertex is in class 0 and the live bitmap `live` is exact for
// the class words as the pass starts (a single-class trim-1 just ran); the
// pass removes pairs from the class words only, so a pair removed meanwhile
// still reads as live, which can only hide a pair, never make one.
template <bool BITS>
__global__ void trim2_kernel(const uint32_t *__restrict__ Fr, const uint32_t *__restrict__ Fc,
                             const uint32_t *__restrict__ Br, const uint32_t *__restrict__ Bc,
                             uint32_t *cls, const uint32_t *__restrict__ live, uint32_t *lab, uint32_t V,
                             uint32_t *ctr)
{
    uint32_t pairs = 0;
    WAVE_LOOP(v, valid, V)
    {
        if (!valid) continue;
        const uint32_t c = BITS ? 0u : cls[v];
        if (BITS ? !same_live<true>(cls, live, v, 0u) : c == DEAD) continue;
        // a single live in-neighbour k that v also points to and whose only
        // live in-neighbour is v, or the same with out-neighbours
        uint32_t k = 0, kk = 0;
        bool hit = live_deg2<BITS>(Br, Bc, cls, live, v, c, &k) == 1 && has_arc(Fr, Fc, v, k) &&
                   live_deg2<BITS>(Br, Bc, cls, live, k, c, &kk) == 1;
        if (!hit) {
            k = 0;
            hit = live_deg2<BITS>(Fr, Fc, cls, live, v, c, &k) == 1 && has_arc(Br, Bc, v, k) &&
                  live_deg2<BITS>(Fr, Fc, cls, live, k, c, &kk) == 1;
        }
        if (hit) {
            uint32_t lo = v < k ? v : k, hi = v < k ? k : v;
            if (atomicCAS(&cls[lo], c, DEAD) == c) {
                lab[lo] = lo | LAB_FINAL;
                lab[hi] = lo | LAB_FINAL;
                __hip_atomic_store(&cls[hi], DEAD, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                pairs++;
            }
        }
    }
    spread_add32(ctr, C_TRIM2, pairs);
}

void k_trim2(hipStream_t s, const DevGraph &g, DevState &st, const uint32_t *live_bits)
{
    if (live_bits)
        trim2_kernel<true><<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(g.Fr, g.Fc, g.Br, g.Bc, st.cls, live_bits,
                                                                     st.lab, g.V, st.ctr);
    else
        trim2_kernel<false><<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(g.Fr, g.Fc, g.Br, g.Bc, st.cls, nullptr,
                                                                      st.lab, g.V, st.ctr);
}

// A search starts with no hub pending and the slots of its queue counts
// clear (bottom-up levels of an earlier search may have left them set).
__device__ __forceinline__ void zero_hub_counts(uint32_t *ctr)
{
    for (int i = 0; i < 3; i++) {
        *reinterpret_cast<unsigned long long *>(ctr + CT(C_BIGF + i)) = 0ull;
        *reinterpret_cast<unsigned long long *>(ctr + CT(C_BIGB + i)) = 0ull;
    }
    for (int j = 0; j < 6 * kSpread; j++) ctr[CSLOT(C_FWCNT, j)] = 0;
}

// ------------------------------------------------- visited bitmaps ---
// Bits of finished vertices (and of slot 0 and the padding) start set in
// the visited-or-finished words of both bitmaps; the finished words beside
// them keep the finished set alone so a bottom-up pass can tell "visited"
// from "finished".  Counts the live vertices.
//
// With ELECT (a round where every live vertex is in one class) the same
// sweep elects the first pivot: the largest pivot key, whose low word names
// its vertex, so no second sweep is needed to find it.  Each workgroup
// leaves its count and key in `bred`; vis_finish_kernel reduces them with
// one workgroup (thousands of same-address atomics would serialise).
constexpr uint32_t VIS_CHUNKS = 8;

template <bool ELECT>
__global__ __launch_bounds__(256) void init_vis_kernel(const uint32_t *__restrict__ cls,
                                                       const uint32_t *__restrict__ live_bits,
                                                       const uint32_t *__restrict__ DP, uint32_t *fvis,
                                                       uint32_t *bvis, uint64_t *bred, uint32_t V,
                                                       uint32_t nchunks)
{
    __shared__ uint64_t s_best[4];
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    uint32_t live = 0;
    uint64_t best = 0;
    // VIS_CHUNKS chunks per step keep that many loads in flight per wave
    for (uint32_t ch0 = wave; ch0 < nchunks; ch0 += VIS_CHUNKS * nwaves) {
        // every chunk's loads are issued before any is used, without
        // per-lane branches (a load and its use in one branch waited for
        // each in turn): a lane past the end reads vertex 0's entries
        uint32_t c[VIS_CHUNKS], dp[VIS_CHUNKS], x[VIS_CHUNKS];
        bool in[VIS_CHUNKS];
#pragma unroll
        for (uint32_t h = 0; h < VIS_CHUNKS; h++) {
            const uint32_t ch = ch0 + h * nwaves, v = ch * 64 + lane();
            in[h] = ch < nchunks && v <= V;
            const uint32_t vv = in[h] ? v : 0u;
            // one class: the live bitmap (2 words a chunk instead of 64)
            x[h] = live_bits ? live_bits[vv >> 5] : cls[vv];
            if (ELECT) dp[h] = DP[vv];
        }
#pragma unroll
        for (uint32_t h = 0; h < VIS_CHUNKS; h++) {
            const uint32_t v = (ch0 + h * nwaves) * 64 + lane();
            if (live_bits)
                c[h] = in[h] && ((x[h] >> (v & 31)) & 1u) ? 0u : DEAD;
            else
                c[h] = in[h] ? x[h] : DEAD;
            if (ELECT && !in[h]) dp[h] = 0u;
        }
#pragma unroll
        for (uint32_t h = 0; h < VIS_CHUNKS; h++) {
            const uint32_t ch = ch0 + h * nwaves, v = ch * 64 + lane();
            const bool dead = c[h] == DEAD;
            live += !dead;
            if (ELECT && !dead) {
                const uint64_t k = ((uint64_t)dp[h] << 32) | (uint64_t)(0xFFFFFFFFu - v);   // pivot_key
                best = k > best ? k : best;
            }
            const uint64_t m = __ballot(dead);
            // the chunk's two groups, each as (visited-or-finished, finished)
            const uint32_t w = (lane() & 2) ? (uint32_t)(m >> 32) : (uint32_t)m;
            uint32_t *dst = lane() < 4 ? fvis : bvis;
            if (lane() < 8 && ch < nchunks) dst[4 * ch + (lane() & 3)] = w;
        }
    }
    if (ELECT) {
        for (int o = 32; o > 0; o >>= 1) {
            const uint64_t x = __shfl_xor(best, o);
            best = x > best ? x : best;
        }
        if (lane() == 0) s_best[threadIdx.x >> 6] = best;
    }
    const uint64_t t = block_sum(live);   // synchronises the workgroup
    if (threadIdx.x == 0) {
        bred[2 * blockIdx.x] = t;
        if (ELECT) {
            for (uint32_t w = 1; w < 4; w++) best = s_best[w] > best ? s_best[w] : best;
            bred[2 * blockIdx.x + 1] = best;
        }
    }
}

// Sums the live counts into C_NLIVE; with ELECT also stores the winning key
// in piv[0] and seeds its vertex (if any) in both directions, like
// select_kernel + seed_counts_kernel, and writes the level-0 state of a
// direction-optimised search in both directions (what do_begin_kernel
// writes), so that search needs no start-up launch.
template <bool ELECT>
__global__ __launch_bounds__(1024) void vis_finish_kernel(const uint64_t *__restrict__ bred, uint32_t nblocks,
                                                          uint32_t *ctr, uint64_t *piv, uint32_t *fvis,
                                                          uint32_t *bvis, uint32_t *qf, uint32_t *qb,
                                                          uint32_t *pivlist, uint32_t *aux)
{
    __shared__ uint64_t s_best[16];
    uint64_t live = 0, best = 0;
    for (uint32_t i = threadIdx.x; i < nblocks; i += blockDim.x) {
        live += bred[2 * i];
        if (ELECT) best = bred[2 * i + 1] > best ? bred[2 * i + 1] : best;
    }
    if (ELECT) {
        for (int o = 32; o > 0; o >>= 1) {
            const uint64_t x = __shfl_xor(best, o);
            best = x > best ? x : best;
        }
        if (lane() == 0) s_best[threadIdx.x >> 6] = best;
    }
    const uint64_t t = block_sum(live);
    if (threadIdx.x != 0) return;
    ctr[CT(C_NLIVE)] = (uint32_t)t;
    if (!ELECT) return;
    for (uint32_t w = 1; w < (blockDim.x >> 6); w++) best = s_best[w] > best ? s_best[w] : best;
    piv[0] = best;
    const uint32_t n = best ? 1u : 0u;
    const uint32_t v = key_vertex(best);
    if (n) {
        pivlist[0] = v;
        qf[0] = v;
        qb[0] = v;
        fvis[vword(v)] |= 1u << (v & 31);
        bvis[vword(v)] |= 1u << (v & 31);
        aux[v] = DEAD;   // the update takes the minimum of the pivot's SCC here
    }
    ctr[CT(C_NPIV)] = n;
    ctr[CT(C_PIV0)] = n ? v : 0u;
    for (int i = 0; i < 3; i++) {
        ctr[CT(C_FWCNT + i)] = i == 0 ? n : 0u;
        ctr[CT(C_BWCNT + i)] = i == 0 ? n : 0u;
    }
    zero_hub_counts(ctr);
    DirState ds;
    ds.mode = DIR_TD;
    ds.end = 0;
    ds.nf = n;
    ds.pad = 0;
    ds.visited = n;
    store_dir_state(ctr, 0, 0, ds);
    store_dir_state(ctr, 0, 1, ds);
}

void k_init_vis(hipStream_t s, const DevGraph &g, DevState &st, bool elect_first, const uint32_t *live_bits)
{
    uint32_t nchunks = st.nwords / 2;
    uint32_t blocks = (nchunks + 3) / 4;
    blocks = blocks < 1 ? 1 : (blocks > VIS_BLOCKS ? VIS_BLOCKS : blocks);
    if (elect_first) {
        init_vis_kernel<true><<<blocks, SWEEP_BLOCK, 0, s>>>(st.cls, live_bits, g.DP, st.fvis, st.bvis,
                                                             st.bred, g.V, nchunks);
        vis_finish_kernel<true><<<1, 1024, 0, s>>>(st.bred, blocks, st.ctr, st.piv, st.fvis, st.bvis, st.qf[0],
                                                   st.qb[0], st.pivlist, st.aux);
    } else {
        init_vis_kernel<false><<<blocks, SWEEP_BLOCK, 0, s>>>(st.cls, live_bits, g.DP, st.fvis, st.bvis,
                                                              st.bred, g.V, nchunks);
        vis_finish_kernel<false><<<1, 1024, 0, s>>>(st.bred, blocks, st.ctr, st.piv, st.fvis, st.bvis, st.qf[0],
                                                    st.qb[0], st.pivlist, st.aux);
    }
}

// ------------------------------------------------------------ pivots ---
// pollForPivots / pollForFirstPivot (scc_kernels.cu:57-74, 643-656) with a
// 64-bit key so the winner is the true argmax and every class elects one
// vertex (defects D4, D10).  A wave whose live lanes share a class does one
// atomic.  The key's degree product is DP[v] (built for every live vertex
// at the start of the decomposition), read with the class word rather than
// four row offsets after it.  Each chunk reads its classes' current maxima
// just before its atomics, so most lanes see a maximum an earlier chunk has
// raised and issue none (deciding four chunks against one read of the
// maxima, with their atomics issued together, doubled the time on a graph
// with about 600 classes: more lanes act on maxima that are out of date).
__global__ void elect_kernel(const uint32_t *__restrict__ DP, const uint32_t *__restrict__ cls,
                             unsigned long long *piv, uint32_t V)
{
    WAVE_LOOP(v, valid, V)
    {
        const uint32_t vv = valid ? v : 0u;   // vertex 0: the dummy, dead
        const uint32_t c = cls[vv], dp = DP[vv];
        const bool ok = c != DEAD;
        const uint64_t live = __ballot(ok);
        if (!live) continue;
        const int L = __ffsll((unsigned long long)live) - 1;
        const uint32_t c0 = __shfl(c, L);
        const unsigned long long key = ok ? (((uint64_t)dp << 32) | (uint64_t)(0xFFFFFFFFu - v)) : 0ull;
        if (__all(!ok || c == c0)) {
            unsigned long long m = key;
            for (int o = 32; o > 0; o >>= 1) {
                unsigned long long x = __shfl_xor(m, o);
                m = x > m ? x : m;
            }
            if ((int)lane() == L && m > piv[c0]) atomicMax(&piv[c0], m);
        } else if (ok && key > piv[c]) {
            atomicMax(&piv[c], key);
        }
    }
}

// First pivot: every live vertex is in class 0, so each workgroup reduces
// its maximum key first and issues one atomicMax.
void k_elect(hipStream_t s, const DevGraph &g, DevState &st)
{
    elect_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(g.DP, st.cls,
                                                           reinterpret_cast<unsigned long long *>(st.piv), g.V);
}

// selectPivots (scc_kernels.cu:4-19): the winner of each class seeds the
// FW (and BW) frontier.
__global__ void select_kernel(const uint32_t *__restrict__ Fr, const uint32_t *__restrict__ Br,
                              const uint32_t *__restrict__ cls, const uint64_t *__restrict__ piv,
                              uint32_t *fvis, uint32_t *bvis, uint32_t *qf, uint32_t *qb,
                              uint32_t *pivlist, uint32_t *ctr, uint32_t V, int seed_bw)
{
    WAVE_LOOP(v, valid, V)
    {
        if (!valid) continue;
        const uint32_t c = cls[v];
        if (c == DEAD || piv[c] != pivot_key(Fr, Br, v)) continue;
        uint32_t i = atomicAdd(&ctr[CT(C_NPIV)], 1u);
        pivlist[i] = v;
        qf[i] = v;
        atomicOr(&fvis[vword(v)], 1u << (v & 31));
        if (seed_bw) {
            qb[i] = v;
            atomicOr(&bvis[vword(v)], 1u << (v & 31));
        }
    }
}

// Queue counts for level 0 and the two counters after it.
__global__ void seed_counts_kernel(uint32_t *ctr, const uint32_t *pivlist, int seed_bw)
{
    uint32_t n = ctr[CT(C_NPIV)];
    ctr[CT(C_PIV0)] = n ? pivlist[0] : 0u;
    ctr[CT(C_FWCNT + 0)] = n;
    ctr[CT(C_FWCNT + 1)] = 0;
    ctr[CT(C_FWCNT + 2)] = 0;
    ctr[CT(C_BWCNT + 0)] = seed_bw ? n : 0;
    ctr[CT(C_BWCNT + 1)] = 0;
    ctr[CT(C_BWCNT + 2)] = 0;
    zero_hub_counts(ctr);
}

void k_select(hipStream_t s, const DevGraph &g, DevState &st, int seed_bw)
{
    select_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(g.Fr, g.Br, st.cls, st.piv, st.fvis,
                                                            st.bvis, st.qf[0], st.qb[0],
                                                            st.pivlist, st.ctr, g.V, seed_bw);
    seed_counts_kernel<<<1, 1, 0, s>>>(st.ctr, st.pivlist, seed_bw);
}

// A single source for the phase entry point: vertex src seeds direction dir.
__global__ void seed_single_kernel(uint32_t *vis, uint32_t *q, uint32_t *ctr, uint32_t src, int dir)
{
    q[0] = src;
    vis[vword(src)] |= 1u << (src & 31);
    for (int i = 0; i < 3; i++) ctr[CT(C_FWCNT + i)] = ctr[CT(C_BWCNT + i)] = 0;
    ctr[CT((dir ? C_BWCNT : C_FWCNT) + 0)] = 1;
    zero_hub_counts(ctr);
}

void k_seed_single(hipStream_t s, DevState &st, uint32_t src, int dir)
{
    seed_single_kernel<<<1, 1, 0, s>>>(dir ? st.bvis : st.fvis, dir ? st.qb[0] : st.qf[0], st.ctr,
                                       src, dir);
}

// ------------------------------------------------------------ update ---
// update (scc_kernels.cu:124-142): FW and BW visited -> the pivot's SCC;
// otherwise the class splits into 3*pivot + (fw + 2*bw), which never wraps
// (defect D9) because each vertex is a pivot at most once.
// Each lane takes UPD_CHUNKS vertices per step: their class words and
// visited words load together, then their pivot keys.
//
// With scc_min (a round with one pivot, the first of a decomposition, whose
// vertex is C_PIV0), the smallest vertex of the pivot's SCC also goes to
// scc_min[pivot] (cleared when the pivot was seeded): the canonical label
// of that SCC, so labelling a decomposition this round finishes needs no
// minimum pass.  The minima are combined per workgroup and added behind a
// coherent read, like the label pass's.
constexpr uint32_t UPD_CHUNKS = 2;

__global__ __launch_bounds__(256) void update_kernel(uint32_t *cls, uint32_t *lab, uint64_t *piv,
                                                     const uint32_t *__restrict__ fvis,
                                                     const uint32_t *__restrict__ bvis, uint32_t *ctr, uint32_t V,
                                                     uint32_t *scc_min, const uint32_t *__restrict__ live_bits,
                                                     int lazy)
{
    // the first round (scc_min) also clears its pivot's next-round entries
    // 3p..3p+2 of the pivot table, which this round does not read (its one
    // class is 0 and p >= 1), so no separate clearing launch follows
    if (scc_min && blockIdx.x == 0 && threadIdx.x < 3) {
        const uint32_t p0 = ctr[CT(C_PIV0)];
        if (p0 >= 1 && p0 <= V) piv[3ull * p0 + threadIdx.x] = 0;
    }
    uint32_t live = 0, found = 0, mmin = DEAD;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base <= V; base += UPD_CHUNKS * stride) {
        uint32_t v[UPD_CHUNKS], c[UPD_CHUNKS], fw[UPD_CHUNKS], bw[UPD_CHUNKS];
        uint64_t key[UPD_CHUNKS];
#pragma unroll
        for (uint32_t h = 0; h < UPD_CHUNKS; h++) {
            v[h] = base + h * stride + lane();   // vertex 0, the dummy, is dead
            const bool in = v[h] <= V;
            if (live_bits)   // one class: the live bitmap instead of the class words
                c[h] = in && ((live_bits[v[h] >> 5] >> (v[h] & 31)) & 1u) ? 0u : DEAD;
            else
                c[h] = in ? cls[v[h]] : DEAD;
            fw[h] = in ? fvis[vword(v[h])] : 0u;
            bw[h] = in ? bvis[vword(v[h])] : 0u;
        }
#pragma unroll
        for (uint32_t h = 0; h < UPD_CHUNKS; h++) key[h] = c[h] != DEAD ? piv[c[h]] : 0ull;
#pragma unroll
        for (uint32_t h = 0; h < UPD_CHUNKS; h++) {
            if (c[h] == DEAD) continue;
            uint32_t p = key_vertex(key[h]);
            if (p < 1 || p > V) {          // no election for this class: cannot happen
                atomicAdd(&ctr[CT(C_ANOM)], 1u);
                p = v[h];
            }
            const uint32_t f = (fw[h] >> (v[h] & 31)) & 1u, b = (bw[h] >> (v[h] & 31)) & 1u;
            if (f && b) {
                if (!lazy) {   // lazy: label_from_bits_kernel or label_fixup_kernel writes these
                    lab[v[h]] = p;
                    cls[v[h]] = DEAD;
                }
                found += v[h] == p;
                mmin = v[h] < mmin ? v[h] : mmin;
            } else {
                cls[v[h]] = 3u * p + f + 2u * b;
                live++;
            }
        }
    }
    spread_add32(ctr, C_LIVE, live);
    spread_add32(ctr, C_PIVOT, found);
    if (!scc_min) return;
    __shared__ uint32_t s_min[4];
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t x = __shfl_xor(mmin, o);
        mmin = x < mmin ? x : mmin;
    }
    if (lane() == 0) s_min[threadIdx.x >> 6] = mmin;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = DEAD;
        for (uint32_t w = 0; w < (blockDim.x >> 6); w++) m = s_min[w] < m ? s_min[w] : m;
        const uint32_t p0 = ctr[CT(C_PIV0)];
        if (m != DEAD && p0 >= 1 && p0 <= V &&
            m < __hip_atomic_load(&scc_min[p0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMin(&scc_min[p0], m);
    }
}

void k_update(hipStream_t s, const DevGraph &g, DevState &st, bool scc_min, const uint32_t *live_bits, bool lazy)
{
    update_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(st.cls, st.lab, st.piv, st.fvis,
                                                            st.bvis, st.ctr, g.V, scc_min ? st.aux : nullptr,
                                                            live_bits, lazy ? 1 : 0);
}

// The lazy first round's update a 32-vertex group per lane: with one class
// (0) and one pivot p0 = C_PIV0, the pivot's SCC is the live vertices
// visited both ways (member = live & fw & bw, bitwise), every other live
// vertex gets class 3*p0 + fw + 2*bw (the only per-vertex writes), and the
// counts and the SCC's smallest vertex come from popcounts and the lowest
// set bit.  Members and trimmed vertices keep their words, as in the lazy
// update_kernel.  Also clears p0's next-round pivot entries.
__global__ __launch_bounds__(256) void update_words_kernel(uint32_t *cls, uint64_t *piv,
                                                           const uint32_t *__restrict__ fvis,
                                                           const uint32_t *__restrict__ bvis,
                                                           const uint32_t *__restrict__ live_bits, uint32_t *ctr,
                                                           uint32_t V, uint32_t *scc_min)
{
    const uint32_t p0 = ctr[CT(C_PIV0)];
    if (blockIdx.x == 0 && threadIdx.x < 3 && p0 >= 1 && p0 <= V) piv[3ull * p0 + threadIdx.x] = 0;
    uint32_t live = 0, found = 0, mmin = DEAD;
    const uint32_t ngroups = (V >> 5) + 1;
    for (uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x; gi < ngroups; gi += gridDim.x * blockDim.x) {
        // the group's three words in flight together (the visited pairs are
        // read even for a group with no live vertex, which then adds and
        // writes nothing, rather than after its live word)
        const uint32_t lw = live_bits[gi];
        const uint64_t f = reinterpret_cast<const uint64_t *>(fvis)[gi];
        const uint64_t b = reinterpret_cast<const uint64_t *>(bvis)[gi];
        const uint32_t fv = (uint32_t)f & lw, bv = (uint32_t)b & lw;   // live vertices are not finished
        const uint32_t member = fv & bv;
        uint32_t rest = lw & ~member;
        live += (uint32_t)__popc(rest);
        if (member) {
            const uint32_t m = (gi << 5) + (uint32_t)(__ffs(member) - 1);
            mmin = m < mmin ? m : mmin;
            if ((p0 >> 5) == gi) found += (member >> (p0 & 31)) & 1u;
        }
        while (rest) {
            const uint32_t bit = (uint32_t)(__ffs(rest) - 1);
            rest &= rest - 1;
            cls[(gi << 5) + bit] = 3u * p0 + ((fv >> bit) & 1u) + 2u * ((bv >> bit) & 1u);
        }
    }
    spread_add32(ctr, C_LIVE, live);
    spread_add32(ctr, C_PIVOT, found);
    __shared__ uint32_t s_min[4];
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t x = __shfl_xor(mmin, o);
        mmin = x < mmin ? x : mmin;
    }
    if (lane() == 0) s_min[threadIdx.x >> 6] = mmin;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = DEAD;
        for (uint32_t w = 0; w < (blockDim.x >> 6); w++) m = s_min[w] < m ? s_min[w] : m;
        if (m != DEAD && p0 >= 1 && p0 <= V &&
            m < __hip_atomic_load(&scc_min[p0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMin(&scc_min[p0], m);
    }
}

void k_update_words(hipStream_t s, const DevGraph &g, DevState &st, const uint32_t *live_bits)
{
    const uint64_t ngroups = ((uint64_t)g.V >> 5) + 1;
    uint64_t b = (ngroups + SWEEP_BLOCK - 1) / SWEEP_BLOCK;
    b = b < 1 ? 1 : (b > 4096 ? 4096 : b);   // a group per lane up to 2^25 vertices
    update_words_kernel<<<(uint32_t)b, SWEEP_BLOCK, 0, s>>>(st.cls, st.piv, st.fvis, st.bvis, live_bits, st.ctr,
                                                            g.V, st.aux);
}

// Members of the first pivot's SCC after a lazy first round: visited by
// both searches and not finished when the bitmaps were initialised (each
// bitmap's second word per group holds the finished bits).
__device__ __forceinline__ bool first_member(const uint32_t *fvis, const uint32_t *bvis, uint32_t v)
{
    const uint64_t f = reinterpret_cast<const uint64_t *>(fvis)[v >> 5];
    const uint64_t b = reinterpret_cast<const uint64_t *>(bvis)[v >> 5];
    const uint32_t fv = (uint32_t)f & ~(uint32_t)(f >> 32), bv = (uint32_t)b & ~(uint32_t)(b >> 32);
    return ((fv & bv) >> (v & 31)) & 1u;
}

// A lazy first round that left live vertices: the labels and the members'
// class words the eager round would have written (members: the pivot,
// non-final; every other vertex: itself, final), before the later phases.
// Trimmed vertices (no live bit) get class DEAD too; live non-members keep
// the class the update gave them.
__global__ void label_fixup_kernel(uint32_t *cls, uint32_t *lab, const uint32_t *__restrict__ fvis,
                                   const uint32_t *__restrict__ bvis, const uint32_t *__restrict__ live_bits,
                                   const uint32_t *__restrict__ ctr, uint32_t V)
{
    const uint32_t p0 = ctr[CT(C_PIV0)];
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x + 1; v <= V; v += gridDim.x * blockDim.x) {
        if (first_member(fvis, bvis, v)) {
            lab[v] = p0;
            cls[v] = DEAD;
        } else {
            lab[v] = v | LAB_FINAL;
            if (!((live_bits[v >> 5] >> (v & 31)) & 1u)) cls[v] = DEAD;
        }
    }
}

void k_label_fixup(hipStream_t s, const DevGraph &g, DevState &st, const uint32_t *live_bits)
{
    label_fixup_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(st.cls, st.lab, st.fvis, st.bvis, live_bits,
                                                                 st.ctr, g.V);
}

// The next round's classes are 3p+k for this round's pivots p: clear them.
__global__ void clear_piv_kernel(const uint32_t *__restrict__ pivlist, uint64_t *piv,
                                 const uint32_t *ctr)
{
    const uint32_t n = ctr[CT(C_NPIV)];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t p = pivlist[i];
        piv[3ull * p] = 0;
        piv[3ull * p + 1] = 0;
        piv[3ull * p + 2] = 0;
    }
}

void k_clear_piv(hipStream_t s, DevState &st, uint32_t max_pivots)
{
    uint32_t b = (max_pivots + SWEEP_BLOCK - 1) / SWEEP_BLOCK;
    b = b < 1 ? 1 : (b > 1024 ? 1024 : b);
    clear_piv_kernel<<<b, SWEEP_BLOCK, 0, s>>>(st.pivlist, st.piv, st.ctr);
}

// |R| and |E_R| of the first FW and BW searches (before update).
__global__ void reach_stats_kernel(const uint32_t *__restrict__ Fr, const uint32_t *__restrict__ Br,
                                   const uint32_t *__restrict__ cls,
                                   const uint32_t *__restrict__ fvis,
                                   const uint32_t *__restrict__ bvis, uint32_t *ctr, uint32_t V)
{
    uint64_t fv = 0, fe = 0, bv = 0, be = 0;
    WAVE_LOOP(v, valid, V)
    {
        if (!valid || cls[v] == DEAD) continue;
        if (bit_of(fvis, v)) { fv++; fe += Fr[v + 1] - Fr[v]; }
        if (bit_of(bvis, v)) { bv++; be += Br[v + 1] - Br[v]; }
    }
    wave_add64(&ctr[CT(C_FWV)], fv);
    wave_add64(&ctr[CT(C_FWE)], fe);
    wave_add64(&ctr[CT(C_BWV)], bv);
    wave_add64(&ctr[CT(C_BWE)], be);
}

void k_reach_stats(hipStream_t s, const DevGraph &g, DevState &st)
{
    reach_stats_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(g.Fr, g.Br, st.cls, st.fvis,
                                                                 st.bvis, st.ctr, g.V);
}

// -------------------------------------------------- "WCC" min label ---
// assignUniqueRange + propagateRange1/2 (scc_kernels.cu:477-536): label =
// min id reachable over live out-arcs; pointer jumping follows label
// chains.  DEAD is the largest value, so finished vertices never win a min.
__global__ void minlabel_init_kernel(uint32_t *cls, uint32_t V)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x + 1; v <= V; v += gridDim.x * blockDim.x)
        if (cls[v] != DEAD) cls[v] = v;
}

void k_minlabel_init(hipStream_t s, const DevGraph &g, DevState &st)
{
    minlabel_init_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(st.cls, g.V);
}

// m folded (minimum, or maximum with MAX) with val[w] over the arcs
// [b, e) of C: K arcs per step, all K columns and then all K values in
// flight before any is used (a step past e repeats the last arc, which
// leaves a minimum or maximum unchanged).
template <bool MAX, uint32_t K>
__device__ __forceinline__ uint32_t fold_arcs(const uint32_t *__restrict__ C, const uint32_t *val, uint32_t b,
                                              uint32_t e, uint32_t m)
{
    for (uint32_t i = b; i < e; i += K) {
        uint32_t w[K], x[K];
#pragma unroll
        for (uint32_t k = 0; k < K; k++) w[k] = C[min(i + k, e - 1)];
#pragma unroll
        for (uint32_t k = 0; k < K; k++) x[k] = val[w[k]];
#pragma unroll
        for (uint32_t k = 0; k < K; k++) m = MAX ? (x[k] > m ? x[k] : m) : (x[k] < m ? x[k] : m);
    }
    return m;
}
constexpr uint32_t MINLABEL_K = 4;   // arcs per step of the min-label propagation
constexpr uint32_t COLOR_K = 4;      // and of the colour propagation

// At most 80 scalar registers, like the search kernel (87 without the cap):
// the 4096^2 grid's min-label rounds 0.50 -> 0.40 ms, the untrimmed R-MAT
// scale 22 colouring rounds 1.10 -> 1.06 ms.
__global__ __attribute__((amdgpu_num_sgpr(80))) void minlabel_prop_kernel(const uint32_t *__restrict__ Fr,
                                     const uint32_t *__restrict__ Fc, uint32_t *cls, uint32_t *ctr,
                                     uint32_t V)
{
    uint32_t changed = 0;
    WAVE_LOOP(v, valid, V)
    {
        // label and row offsets together (past V, vertex 0's: the dummy,
        // dead); an inactive lane folds an empty range, so the offsets are
        // used on every path and are not loaded after the label test
        const uint32_t vv = valid ? v : 0u;
        const uint32_t m0 = cls[vv], b = Fr[vv], e = Fr[vv + 1];
        const bool act = valid && m0 != DEAD;
        uint32_t m = fold_arcs<false, MINLABEL_K>(Fc, cls, b, act ? e : b, m0);
        if (!act) continue;
        for (int it = 0; it < 32; it++) {
            uint32_t x = cls[m];
            if (x >= m) break;
            m = x;
        }
        if (m < m0) {
            __hip_atomic_store(&cls[v], m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            changed = 1;
        }
    }
    spread_add32(ctr, C_CHANGED, changed);
}

void k_minlabel_prop(hipStream_t s, const DevGraph &g, DevState &st)
{
    minlabel_prop_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(g.Fr, g.Fc, st.cls, st.ctr, g.V);
}

// class = 3 * label; the class representative clears its pivot slot.
__global__ void minlabel_finish_kernel(uint32_t *cls, uint64_t *piv, uint32_t V)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x + 1; v <= V; v += gridDim.x * blockDim.x) {
        uint32_t m = cls[v];
        if (m == DEAD) continue;
        if (m == v) piv[3ull * v] = 0;
        cls[v] = 3u * m;
    }
}

void k_minlabel_finish(hipStream_t s, const DevGraph &g, DevState &st)
{
    minlabel_finish_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(st.cls, st.piv, g.V);
}

// ----------------------------------------------------- colouring ---
// colorPropagation (scc_kernels.cu:539-564): colour = max id reachable over
// live out-arcs, with pointer jumping added.  aux[v] = 0 for finished v.
__global__ void color_init_kernel(const uint32_t *__restrict__ cls, uint32_t *aux, uint32_t V)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= V; v += gridDim.x * blockDim.x)
        aux[v] = (v && cls[v] != DEAD) ? v : 0u;
}

void k_color_init(hipStream_t s, const DevGraph &g, DevState &st)
{
    color_init_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(st.cls, st.aux, g.V);
}

__global__ __attribute__((amdgpu_num_sgpr(80))) void color_prop_kernel(const uint32_t *__restrict__ Fr,
                                                                     const uint32_t *__restrict__ Fc,
                                                                     uint32_t *aux, uint32_t *ctr, uint32_t V)
{
    uint32_t changed = 0;
    WAVE_LOOP(v, valid, V)
    {
        // colour and row offsets together (past V, vertex 0's: colour 0);
        // an inactive lane folds an empty range, as in minlabel_prop_kernel
        const uint32_t vv = valid ? v : 0u;
        const uint32_t m0 = aux[vv], b = Fr[vv], e = Fr[vv + 1];
        const bool act = valid && m0 != 0;
        uint32_t m = fold_arcs<true, COLOR_K>(Fc, aux, b, act ? e : b, m0);
        if (!act) continue;
        for (int it = 0; it < 32; it++) {
            uint32_t x = aux[m];
            if (x <= m) break;
            m = x;
        }
        if (m > m0) {
            __hip_atomic_store(&aux[v], m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            changed = 1;
        }
    }
    spread_add32(ctr, C_CHANGED, changed);
}

void k_color_prop(hipStream_t s, const DevGraph &g, DevState &st)
{
    color_prop_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(g.Fr, g.Fc, st.aux, st.ctr, g.V);
}

// selectPivotColoring (scc_kernels.cu:567-580): class = colour; a vertex
// whose colour is its own id is a root and seeds the FW frontier.
__global__ void color_roots_kernel(uint32_t *cls, const uint32_t *__restrict__ aux, uint32_t *fvis,
                                   uint32_t *qf, uint32_t *pivlist, uint32_t *ctr, uint32_t V)
{
    WAVE_LOOP(v, valid, V)
    {
        if (!valid || cls[v] == DEAD) continue;
        const uint32_t c = aux[v];
        cls[v] = c;
        if (c == v) {
            uint32_t i = atomicAdd(&ctr[CT(C_NPIV)], 1u);
            pivlist[i] = v;
            qf[i] = v;
            atomicOr(&fvis[vword(v)], 1u << (v & 31));
        }
    }
}

void k_color_roots(hipStream_t s, const DevGraph &g, DevState &st)
{
    color_roots_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(st.cls, st.aux, st.fvis, st.qf[0],
                                                                 st.pivlist, st.ctr, g.V);
    seed_counts_kernel<<<1, 1, 0, s>>>(st.ctr, st.pivlist, 0);
}

// fwdColoring marks each vertex it expands as finished (scc_kernels.cu:607):
// the vertices of colour r reached from root r form SCC(r).
__global__ void color_finish_kernel(uint32_t *cls, uint32_t *lab, const uint32_t *__restrict__ fvis,
                                    uint32_t *ctr, uint32_t V)
{
    uint32_t live = 0, found = 0;
    WAVE_LOOP(v, valid, V)
    {
        if (!valid) continue;
        const uint32_t c = cls[v];
        if (c == DEAD) continue;
        if (bit_of(fvis, v)) {
            lab[v] = c;
            cls[v] = DEAD;
            found += c == v;
        } else {
            live++;
        }
    }
    spread_add32(ctr, C_LIVE, live);
    spread_add32(ctr, C_PIVOT, found);
}

void k_color_finish(hipStream_t s, const DevGraph &g, DevState &st)
{
    color_finish_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(st.cls, st.lab, st.fvis, st.ctr,
                                                                  g.V);
}

// ------------------------------------------------ canonical labels ---
// lab[v] -> min vertex id with the same representative; count roots and
// the largest SCC.  A wave's lanes that share the first lane's label
// combine into one atomic (the giant SCC would otherwise serialise).
// With a gate (the counters), a label pass queued before the host knows
// whether any vertex is still live is a no-op unless C_LIVE (its own word
// plus its slots, summed by each wave) is 0.
__device__ __forceinline__ bool gate_closed(const uint32_t *gate)
{
    if (!gate) return false;
    const uint32_t l = lane();
    uint32_t x = l == 0 ? gate[CT(C_LIVE)] : (l <= (uint32_t)kSpread ? gate[CSLOT(C_LIVE, l - 1)] : 0u);
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x != 0;
}

__global__ void canon_reset_kernel(uint32_t *aux, uint32_t V, const uint32_t *gate)
{
    if (gate_closed(gate)) return;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= V; v += gridDim.x * blockDim.x) aux[v] = DEAD;
}

// Lanes of a wave that share a label are combined before touching memory:
// up to MATCH_ROUNDS times the lowest pending lane leads and every lane with
// its label joins; what is left (a wave with many distinct labels) goes
// lane by lane.  A giant SCC's members then cost one atomic per wave
// instead of one per vertex.
constexpr int MATCH_ROUNDS = 8;

// Each lane takes CANON_CHUNKS vertices per step, one per 64-vertex chunk
// of its wave, with all their label loads in flight together (then, in the
// apply pass, all their representative loads): one load per step leaves
// these passes waiting on memory latency.
constexpr uint32_t CANON_CHUNKS = 8;

// Like the size counts below, each wave keeps a running minimum for one
// label several of its lanes shared and writes it once at the end.
__global__ __launch_bounds__(256) void canon_min_kernel(uint32_t *lab, uint32_t *aux, uint32_t *ctr, uint32_t V,
                                                        const uint32_t *gate)
{
    if (gate_closed(gate)) return;
    uint32_t heavy = DEAD, hmin = DEAD;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < V; base += CANON_CHUNKS * stride) {
        uint32_t vs[CANON_CHUNKS], rs[CANON_CHUNKS];
#pragma unroll
        for (uint32_t h = 0; h < CANON_CHUNKS; h++) {
            vs[h] = base + h * stride + lane() + 1;
            rs[h] = vs[h] <= V ? lab[vs[h]] : DEAD;
        }
#pragma unroll
        for (uint32_t h = 0; h < CANON_CHUNKS; h++) {
            const uint32_t v = vs[h];
            const bool valid = v <= V;
            uint32_t r = rs[h];
            const bool final_lab = valid && (r & LAB_FINAL);
            const uint32_t rr = r & ~LAB_FINAL;
            if (valid && (rr < 1 || rr > V || (final_lab && rr > v))) {
                // every vertex must hold a representative (a final one is <= v)
                atomicAdd(&ctr[CT(C_ANOM)], 1u);
                lab[v] = r = v | LAB_FINAL;
            }
            bool pending = valid && !(r & LAB_FINAL);
            for (int it = 0; it < MATCH_ROUNDS; it++) {
                const uint64_t pm = __ballot(pending);
                if (!pm) break;
                const int L = __ffsll((unsigned long long)pm) - 1;
                const uint32_t rl = __shfl(r, L);
                const bool mine = pending && r == rl;
                const uint64_t mm = __ballot(mine);
                uint32_t m = mine ? v : DEAD;
                for (int o = 32; o > 0; o >>= 1) {
                    uint32_t x = __shfl_xor(m, o);
                    m = x < m ? x : m;
                }
                if (rl == heavy) {
                    hmin = m < hmin ? m : hmin;
                } else if (heavy == DEAD && __popcll(mm) > 1) {
                    heavy = rl;
                    hmin = m;
                } else if ((int)lane() == L && m < aux[rl]) {
                    atomicMin(&aux[rl], m);
                }
                pending = pending && !mine;
            }
            if (pending && v < aux[r]) atomicMin(&aux[r], v);
        }
    }
    // The waves' running minima (every wave of a giant SCC's decomposition
    // holds one for the same label) are combined per workgroup, and one
    // lane adds it only if a coherent read shows it still lowers the value:
    // a plain read sees a stale cached copy, and thousands of waves ending
    // together with an atomic each on one address serialise.
    __shared__ uint32_t s_heavy[4], s_hmin[4];
    const uint32_t wv = threadIdx.x >> 6;
    if (lane() == 0) {
        s_heavy[wv] = heavy;
        s_hmin[wv] = hmin;
    }
    __syncthreads();
    if (lane() == 0 && heavy != DEAD) {
        uint32_t m = hmin;
        bool lead = true;   // the lowest wave holding this label adds for all of them
        for (uint32_t w = 0; w < (blockDim.x >> 6); w++) {
            if (w == wv || s_heavy[w] != heavy) continue;
            if (w < wv) lead = false;
            m = s_hmin[w] < m ? s_hmin[w] : m;
        }
        if (lead && m < __hip_atomic_load(&aux[heavy], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMin(&aux[heavy], m);
    }
}

__global__ __launch_bounds__(256) void canon_apply_kernel(uint32_t *lab, const uint32_t *__restrict__ aux,
                                                          uint32_t *ctr, uint32_t V, const uint32_t *gate)
{
    if (gate_closed(gate)) return;
    uint32_t roots = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < V; base += CANON_CHUNKS * stride) {
        uint32_t vs[CANON_CHUNKS], rs[CANON_CHUNKS];
#pragma unroll
        for (uint32_t h = 0; h < CANON_CHUNKS; h++) {
            vs[h] = base + h * stride + lane() + 1;
            rs[h] = vs[h] <= V ? lab[vs[h]] : DEAD;
        }
#pragma unroll
        for (uint32_t h = 0; h < CANON_CHUNKS; h++) {
            const uint32_t l = rs[h];
            if (vs[h] <= V) rs[h] = (l & LAB_FINAL) ? (l & ~LAB_FINAL) : aux[l];
        }
#pragma unroll
        for (uint32_t h = 0; h < CANON_CHUNKS; h++) {
            const uint32_t v = vs[h], r = rs[h];
            if (v <= V) {
                lab[v] = r;
                roots += r == v;
            }
        }
    }
    spread_add32(ctr, C_NSCC, roots);
}

// The labels of a decomposition that a lazy first round finished (gated:
// a no-op unless C_LIVE is 0): a member of the pivot's SCC gets the SCC's
// smallest vertex (aux[pivot], left by the update), every other vertex
// (trimmed) itself.  Reads two bitmap words per 32 vertices instead of the
// labels; counts the roots like canon_apply_kernel.
__global__ __launch_bounds__(256) void label_from_bits_kernel(uint32_t *lab, const uint32_t *__restrict__ fvis,
                                                              const uint32_t *__restrict__ bvis,
                                                              const uint32_t *__restrict__ aux, uint32_t *ctr,
                                                              uint32_t V, const uint32_t *gate)
{
    if (gate_closed(gate)) return;
    const uint32_t p0 = ctr[CT(C_PIV0)];
    const uint32_t rmin = p0 >= 1 && p0 <= V ? aux[p0] : DEAD;
    uint32_t roots = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < V; base += CANON_CHUNKS * stride) {
#pragma unroll
        for (uint32_t h = 0; h < CANON_CHUNKS; h++) {
            const uint32_t v = base + h * stride + lane() + 1;
            if (v > V) continue;
            const uint32_t r = first_member(fvis, bvis, v) ? rmin : v;
            lab[v] = r;
            roots += r == v;
        }
    }
    spread_add32(ctr, C_NSCC, roots);
}

void k_label_from_bits(hipStream_t s, const DevGraph &g, DevState &st)
{
    label_from_bits_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(st.lab, st.fvis, st.bvis, st.aux, st.ctr,
                                                                     g.V, st.ctr);
}

void k_canon_apply(hipStream_t s, const DevGraph &g, DevState &st, bool gated)
{
    canon_apply_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(st.lab, st.aux, st.ctr, g.V,
                                                                 gated ? st.ctr : nullptr);
}

void k_canon(hipStream_t s, const DevGraph &g, DevState &st, bool sizes, bool gated)
{
    const uint32_t *gate = gated ? st.ctr : nullptr;
    const uint32_t b = sweep_blocks(g.V);
    canon_reset_kernel<<<b, SWEEP_BLOCK, 0, s>>>(st.aux, g.V, gate);
    canon_min_kernel<<<b, SWEEP_BLOCK, 0, s>>>(st.lab, st.aux, st.ctr, g.V, gate);
    canon_apply_kernel<<<b, SWEEP_BLOCK, 0, s>>>(st.lab, st.aux, st.ctr, g.V, gate);
    if (sizes) k_sizes(s, g, st, gated);
}

// ------------------------------------------------- SCC sizes ---
// From canonical labels (lab[v] = the smallest vertex of v's SCC): cnt[r]
// counts the members of r's SCC other than r itself, so a singleton (most
// SCCs of a power-law graph: trim-1's) costs no atomic at all, and the
// members of a large SCC are combined per wave (MATCH_ROUNDS leaders, then
// one running count per wave for a recurring label) before touching memory.
// The size of the SCC of root r (lab[r] == r) is cnt[r] + 1.  cnt is the
// class-word array, free once the labels are final.
__global__ void size_reset_kernel(uint32_t *cnt, uint32_t V, const uint32_t *gate)
{
    if (gate_closed(gate)) return;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= V; v += gridDim.x * blockDim.x) cnt[v] = 0;
}

__global__ __launch_bounds__(256) void size_count_kernel(const uint32_t *__restrict__ lab, uint32_t *cnt,
                                                         uint32_t V, const uint32_t *gate)
{
    if (gate_closed(gate)) return;
    uint32_t heavy = DEAD, hcount = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < V; base += CANON_CHUNKS * stride) {
        uint32_t vs[CANON_CHUNKS], rs[CANON_CHUNKS];
#pragma unroll
        for (uint32_t h = 0; h < CANON_CHUNKS; h++) {
            vs[h] = base + h * stride + lane() + 1;
            rs[h] = vs[h] <= V ? lab[vs[h]] : 0u;
        }
#pragma unroll
        for (uint32_t h = 0; h < CANON_CHUNKS; h++) {
            const uint32_t r = rs[h];
            bool pending = vs[h] <= V && r != vs[h] && r >= 1 && r <= V;
            for (int it = 0; it < MATCH_ROUNDS; it++) {
                const uint64_t pm = __ballot(pending);
                if (!pm) break;
                const int L = __ffsll((unsigned long long)pm) - 1;
                const uint32_t rl = __shfl(r, L);
                const uint32_t n = (uint32_t)__popcll(__ballot(pending && r == rl));
                if (rl == heavy) {
                    hcount += n;
                } else if (heavy == DEAD && n > 1) {
                    heavy = rl;
                    hcount = n;
                } else if ((int)lane() == L) {
                    atomicAdd(&cnt[rl], n);
                }
                pending = pending && r != rl;
            }
            if (pending) atomicAdd(&cnt[r], 1u);
        }
    }
    if (hcount && lane() == 0) atomicAdd(&cnt[heavy], hcount);
}

// Largest SCC: the largest cnt[r] + 1 over roots r.
__global__ void size_largest_kernel(const uint32_t *__restrict__ lab, const uint32_t *__restrict__ cnt,
                                    uint32_t *ctr, uint32_t V, const uint32_t *gate)
{
    if (gate_closed(gate)) return;
    uint32_t m = 0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x + 1; v <= V; v += gridDim.x * blockDim.x)
        if (lab[v] == v) m = cnt[v] + 1 > m ? cnt[v] + 1 : m;
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t x = __shfl_xor(m, o);
        m = x > m ? x : m;
    }
    if (lane() == 0 && m) atomicMax(&ctr[CT(C_LARGEST)], m);
}

void k_sizes(hipStream_t s, const DevGraph &g, DevState &st, bool gated)
{
    const uint32_t *gate = gated ? st.ctr : nullptr;
    const uint32_t b = sweep_blocks(g.V);
    size_reset_kernel<<<b, SWEEP_BLOCK, 0, s>>>(st.cls, g.V, gate);
    size_count_kernel<<<b, SWEEP_BLOCK, 0, s>>>(st.lab, st.cls, g.V, gate);
    size_largest_kernel<<<b, SWEEP_BLOCK, 0, s>>>(st.lab, st.cls, st.ctr, g.V, gate);
}

// Size histogram from the sizes pass: hist[k] = number of SCCs of k
// vertices (hist has V + 1 words, cleared here), combined per wave like
// the counts (a power-law graph has millions of roots of size 1), then the
// non-zero entries compacted into (size, count) pairs: `pairs` holds up to
// `cap` of them (64-bit words, size then count), *npairs (a counter line)
// the number found.  At most about sqrt(2V) sizes occur, so the list is
// short; the host sorts it.
__global__ __launch_bounds__(256) void hist_count_kernel(const uint32_t *__restrict__ lab,
                                                         const uint32_t *__restrict__ cnt, uint32_t *hist,
                                                         uint32_t V)
{
    uint32_t heavy = 0, hcount = 0;
    for (uint32_t v0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); v0 < V; v0 += gridDim.x * blockDim.x) {
        const uint32_t v = v0 + lane() + 1;
        const uint32_t k = (v <= V && lab[v] == v) ? cnt[v] + 1 : 0u;
        bool pending = k != 0;
        for (int it = 0; it < MATCH_ROUNDS; it++) {
            const uint64_t pm = __ballot(pending);
            if (!pm) break;
            const int L = __ffsll((unsigned long long)pm) - 1;
            const uint32_t kl = __shfl(k, L);
            const uint32_t n = (uint32_t)__popcll(__ballot(pending && k == kl));
            if (kl == heavy) {
                hcount += n;
            } else if (heavy == 0 && n > 1) {
                heavy = kl;
                hcount = n;
            } else if ((int)lane() == L) {
                atomicAdd(&hist[kl], n);
            }
            pending = pending && k != kl;
        }
        if (pending) atomicAdd(&hist[k], 1u);
    }
    if (hcount && lane() == 0) atomicAdd(&hist[heavy], hcount);
}

__global__ void hist_compact_kernel(const uint32_t *__restrict__ hist, uint64_t *pairs, uint32_t cap,
                                    uint32_t *npairs, uint32_t V)
{
    for (uint32_t k0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); k0 < V; k0 += gridDim.x * blockDim.x) {
        const uint32_t k = k0 + lane() + 1;
        const uint32_t c = k <= V ? hist[k] : 0u;
        const uint64_t m = __ballot(c != 0);
        if (!m) continue;
        const int leader = __ffsll((unsigned long long)m) - 1;
        uint32_t base = 0;
        if ((int)lane() == leader) base = atomicAdd(npairs, (uint32_t)__popcll(m));
        base = __shfl(base, leader);
        const uint32_t at = base + (uint32_t)__popcll(m & ((1ull << lane()) - 1));
        if (c && at < cap) {
            pairs[2 * (size_t)at] = k;
            pairs[2 * (size_t)at + 1] = c;
        }
    }
}

void k_size_histogram(hipStream_t s, const DevGraph &g, DevState &st, uint32_t *hist, uint64_t *pairs,
                      uint32_t cap, uint32_t *npairs)
{
    const uint32_t b = sweep_blocks(g.V);
    size_reset_kernel<<<b, SWEEP_BLOCK, 0, s>>>(hist, g.V, nullptr);
    hist_count_kernel<<<b, SWEEP_BLOCK, 0, s>>>(st.lab, st.cls, hist, g.V);
    hist_compact_kernel<<<b, SWEEP_BLOCK, 0, s>>>(hist, pairs, cap, npairs, g.V);
}

// ------------------------------------------------ phase exports ---
__global__ void export_visited_kernel(const uint32_t *__restrict__ cls,
                                      const uint32_t *__restrict__ vis, uint8_t *out, uint32_t V)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= V; v += gridDim.x * blockDim.x)
        out[v] = (v && cls[v] != DEAD && bit_of(vis, v)) ? 1 : 0;
}

void k_export_visited(hipStream_t s, const DevGraph &g, DevState &st, int dir, uint8_t *out_dev)
{
    export_visited_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(st.cls, dir ? st.bvis : st.fvis,
                                                                    out_dev, g.V);
}

__global__ void export_removed_kernel(const uint32_t *__restrict__ cls, uint8_t *out, uint32_t V)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= V; v += gridDim.x * blockDim.x)
        out[v] = (v && cls[v] == DEAD) ? 1 : 0;
}

void k_export_removed(hipStream_t s, const DevGraph &g, DevState &st, uint8_t *out_dev)
{
    export_removed_kernel<<<sweep_blocks(g.V), SWEEP_BLOCK, 0, s>>>(st.cls, out_dev, g.V);
}

} // namespace sccamd
