// kernels_batch.hip -- many independent graphs in one persistent launch.
//
// The reference runs one graph per process with a host round trip after
// every kernel (scc_WCC.cu:683-686).  Here each workgroup pops a graph index
// from a device work queue (largest graphs first) and runs the whole SCC
// pipeline for that graph on its own, with workgroup barriers between the
// phases and every queue count and fixpoint flag in LDS.  Nothing goes back
// to the host until the batch is done.
//
// Phase orders are the reference's: vHong (scc_WCC.cu:573-998) and vSlota
// (scc_Coloring.cu:403-701); the kernels inside each phase mirror the
// single-graph engine (kernels_sweep.hip, kernels_reach.hip) at workgroup
// scope.  Per graph the configuration word selects the algorithm, the trims
// and the lanes per frontier vertex (chosen by the autotuner per size and
// degree bucket).
//
// Layout: graph i's per-vertex arrays start at slot voff[i] (a multiple of
// 64, so bitmap words never straddle two graphs) and span V_i+2 slots; its
// column arrays start at eoff[i]; its pivot table spans 3*(V_i+2) keys from
// 3*voff[i].  Vertex ids inside a graph are local (1..V_i).
#include "batch.h"

namespace sccamd {

// cfg word: bits 0-2 log2(lanes per frontier vertex), bit 3 trim-1, bit 4
// trim-2, bit 5 Slota (colouring) instead of Hong
__host__ __device__ inline uint32_t cfg_lanes(uint32_t c) { return 1u << (c & 7u); }
__host__ __device__ inline bool cfg_t1(uint32_t c) { return (c >> 3) & 1u; }
__host__ __device__ inline bool cfg_t2(uint32_t c) { return (c >> 4) & 1u; }
__host__ __device__ inline bool cfg_slota(uint32_t c) { return (c >> 5) & 1u; }

namespace {

__device__ __forceinline__ uint32_t ln() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint64_t bkey(const uint32_t *Fr, const uint32_t *Br, uint32_t v)
{
    uint64_t p = (uint64_t)(Fr[v + 1] - Fr[v]) * (uint64_t)(Br[v + 1] - Br[v]);
    p = p > 0xFFFFFFFFull ? 0xFFFFFFFFull : p;
    return (p << 32) | (uint64_t)(0xFFFFFFFFu - v);
}

__device__ __forceinline__ uint32_t bkey_vertex(uint64_t k) { return 0xFFFFFFFFu - (uint32_t)k; }

__device__ __forceinline__ bool bbit(const uint32_t *bm, uint32_t v) { return (bm[v >> 5] >> (v & 31)) & 1u; }

// per-workgroup scratch in LDS
struct Lds {
    uint32_t graph;
    uint32_t flag;
    uint32_t cnt[4];          // queue counts: FW in, FW out, BW in, BW out
    uint32_t npiv;
    uint32_t live;
    uint32_t anom;            // a class without a recorded pivot was seen
    uint32_t red[16];
    unsigned long long key;
};

// one graph's views
struct G1 {
    uint32_t V;
    const uint32_t *Fr, *Fc, *Br, *Bc;
    uint32_t *cls, *lab, *aux;
    uint32_t *fvis, *bvis;
    uint64_t *piv;
    uint32_t *q;          // four queues of V+2 entries: FW in/out, BW in/out
    // queue k (0..3) by arithmetic: a runtime index into a pointer array
    // would put the whole struct in scratch memory
    __device__ __forceinline__ uint32_t *queue(uint32_t k) const { return q + (size_t)k * (V + 2); }
    uint32_t *pivlist;
};

template <int BS> __device__ __forceinline__ uint32_t wg_sum(Lds &L, uint32_t x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    __syncthreads();
    if (ln() == 0) L.red[threadIdx.x >> 6] = x;
    __syncthreads();
    uint32_t t = 0;
    for (int w = 0; w < BS / 64; w++) t += L.red[w];
    __syncthreads();
    return t;
}

// LDS-counter push (wave-aggregated)
__device__ __forceinline__ void lpush(bool p, uint32_t v, uint32_t *q, uint32_t *cnt)
{
    const uint64_t act = __ballot(1);
    const uint64_t m = __ballot(p);
    if (!m) return;
    const int leader = __ffsll((unsigned long long)act) - 1;
    uint32_t b = 0;
    if ((int)ln() == leader) b = atomicAdd(cnt, (uint32_t)__popcll(m));
    b = __shfl(b, leader);
    if (p) q[b + (uint32_t)__popcll(m & ((1ull << ln()) - 1))] = v;
}

template <int BS> __device__ void g_init_vis(const G1 &g)
{
    const uint32_t nch = (g.V + 64) / 64;
    for (uint32_t ch = threadIdx.x >> 6; ch < nch; ch += BS / 64) {
        const uint32_t v = ch * 64 + ln();
        const bool dead = v > g.V || g.cls[v] == DEAD;
        const uint64_t m = __ballot(dead);
        if (ln() < 2) g.fvis[2 * ch + ln()] = ln() ? (uint32_t)(m >> 32) : (uint32_t)m;
        else if (ln() < 4) g.bvis[2 * ch + (ln() & 1)] = (ln() & 1) ? (uint32_t)(m >> 32) : (uint32_t)m;
    }
}

// Per-vertex list walks: a step loads GK columns and then issues every
// probe of the step before using any (indexes past the end repeat the last
// arc, which changes no minimum, maximum or existence test).  A walk of
// one arc per step waits for two dependent loads per arc.
constexpr uint32_t GK = 4;

// whether some arc of col[i..e) leads to a vertex of class c
__device__ __forceinline__ bool g_any_cls(const uint32_t *C, uint32_t i, uint32_t e, const uint32_t *cls,
                                          uint32_t c)
{
    bool hit = false;
    for (; !hit && i < e; i += GK) {
        uint32_t w[GK];
#pragma unroll
        for (uint32_t k = 0; k < GK; k++) w[k] = C[min(i + k, e - 1)];
#pragma unroll
        for (uint32_t k = 0; k < GK; k++) hit |= cls[w[k]] == c;
    }
    return hit;
}

// m combined with a[w] over the arcs col[i..e): the largest (MAX) or the smallest
template <bool MAX>
__device__ __forceinline__ uint32_t g_fold(const uint32_t *C, uint32_t i, uint32_t e, const uint32_t *a, uint32_t m)
{
    for (; i < e; i += GK) {
        uint32_t w[GK];
#pragma unroll
        for (uint32_t k = 0; k < GK; k++) w[k] = C[min(i + k, e - 1)];
#pragma unroll
        for (uint32_t k = 0; k < GK; k++) {
            const uint32_t x = a[w[k]];
            m = MAX ? (x > m ? x : m) : (x < m ? x : m);
        }
    }
    return m;
}

// trim-1 (scc_kernels.cu:187-233) to its fixpoint or one pass
template <int BS> __device__ uint32_t g_trim1(Lds &L, const G1 &g, bool fixpoint)
{
    uint32_t total = 0;
    for (uint32_t pass = 0; pass <= g.V; pass++) {   // each pass removes a vertex or ends
        uint32_t removed = 0;
        for (uint32_t v = threadIdx.x + 1; v <= g.V; v += BS) {
            // class and both rows' offsets together; a dead vertex scans
            // empty ranges, so the offsets are used on every path and are
            // not loaded after the class test or the first scan
            const uint32_t c = g.cls[v];
            const uint32_t bb = g.Br[v], be = g.Br[v + 1], fb = g.Fr[v], fe = g.Fr[v + 1];
            const bool act = c != DEAD;
            const bool kin = g_any_cls(g.Bc, bb, act ? be : bb, g.cls, c);
            const bool keep = kin && g_any_cls(g.Fc, fb, fe, g.cls, c);
            if (act && !keep) {
                g.lab[v] = v | LAB_FINAL;
                __hip_atomic_store(&g.cls[v], DEAD, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                removed++;
            }
        }
        const uint32_t n = wg_sum<BS>(L, removed);
        total += n;
        if (!fixpoint || !n) break;
    }
    return total;
}

__device__ __forceinline__ int g_deg2(const uint32_t *R, const uint32_t *C, const uint32_t *cls,
                                      uint32_t v, uint32_t c, uint32_t *k)
{
    int d = 0;
    for (uint32_t i = R[v], e = R[v + 1]; i < e; i++) {
        const uint32_t w = C[i];
        if (cls[w] != c) continue;
        if (++d == 2) return 2;
        *k = w;
    }
    return d;
}

__device__ __forceinline__ bool g_arc(const uint32_t *R, const uint32_t *C, uint32_t v, uint32_t k)
{
    for (uint32_t i = R[v], e = R[v + 1]; i < e; i++)
        if (C[i] == k) return true;
    return false;
}

// trim-2 (scc_kernels.cu:236-357), claim by CAS on the smaller id
template <int BS> __device__ void g_trim2(const G1 &g)
{
    for (uint32_t v = threadIdx.x + 1; v <= g.V; v += BS) {
        const uint32_t c = g.cls[v];
        if (c == DEAD) continue;
        uint32_t k = 0, kk = 0;
        bool hit = g_deg2(g.Br, g.Bc, g.cls, v, c, &k) == 1 && g_arc(g.Fr, g.Fc, v, k) &&
                   g_deg2(g.Br, g.Bc, g.cls, k, c, &kk) == 1;
        if (!hit) {
            k = 0;
            hit = g_deg2(g.Fr, g.Fc, g.cls, v, c, &k) == 1 && g_arc(g.Br, g.Bc, v, k) &&
                  g_deg2(g.Fr, g.Fc, g.cls, k, c, &kk) == 1;
        }
        if (hit) {
            const uint32_t lo = v < k ? v : k, hi = v < k ? k : v;
            if (atomicCAS(&g.cls[lo], c, DEAD) == c) {
                g.lab[lo] = lo | LAB_FINAL;
                g.lab[hi] = lo | LAB_FINAL;
                __hip_atomic_store(&g.cls[hi], DEAD, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
    }
    __syncthreads();
}

// FW and/or BW level loop from queues [0] with counts L.cnt[0] / L.cnt[2]
template <int BS> __device__ void g_reach(Lds &L, const G1 &g, bool do_bw, bool single, uint32_t W)
{
    int in = 0;
    for (uint32_t level = 0; level <= g.V; level++) {   // each level claims a vertex or ends
        const uint32_t nf = L.cnt[0], nb = do_bw ? L.cnt[2] : 0u;
        if (!nf && !nb) break;
        __syncthreads();
        if (threadIdx.x == 0) {
            L.cnt[1] = 0;
            L.cnt[3] = 0;
        }
        __syncthreads();
        const uint32_t grp = threadIdx.x / W, sub = threadIdx.x % W, ngrp = BS / W;
        for (int d = 0; d < 2; d++) {
            const uint32_t n = d ? nb : nf;
            const uint32_t *R = d ? g.Br : g.Fr, *C = d ? g.Bc : g.Fc;
            uint32_t *vis = d ? g.bvis : g.fvis;
            const uint32_t *qin = g.queue(2 * d + in);
            uint32_t *qout = g.queue(2 * d + (in ^ 1));
            uint32_t *cnt = &L.cnt[d ? 3 : 1];
            // two frontier vertices per lane group in flight: their queue
            // entries, rows, columns and probes are loaded side by side
            // (vertex 0, the dummy, stands for "none": its row is empty)
            for (uint32_t i0 = grp; i0 < n; i0 += 2 * ngrp) {
                uint32_t v[2], b[2], e[2], c[2];
#pragma unroll
                for (uint32_t u = 0; u < 2; u++) v[u] = i0 + u * ngrp < n ? qin[i0 + u * ngrp] : 0u;
#pragma unroll
                for (uint32_t u = 0; u < 2; u++) {
                    b[u] = R[v[u]];
                    e[u] = R[v[u] + 1];
                    c[u] = single ? 0u : g.cls[v[u]];
                }
                const uint32_t d0 = e[0] - b[0], d1 = e[1] - b[1];
                const uint32_t dm = d0 > d1 ? d0 : d1;
                for (uint32_t j = sub; j < dm; j += W) {
                    uint32_t w[2] = {0u, 0u};
                    bool ok[2] = {j < d0, j < d1};
#pragma unroll
                    for (uint32_t u = 0; u < 2; u++)
                        if (ok[u]) w[u] = C[b[u] + j];
                    uint32_t word[2] = {~0u, ~0u}, wc[2] = {0u, 0u};
#pragma unroll
                    for (uint32_t u = 0; u < 2; u++)
                        if (ok[u]) {
                            word[u] = vis[w[u] >> 5];
                            if (!single) wc[u] = g.cls[w[u]];
                        }
                    bool claimed[2] = {false, false};
#pragma unroll
                    for (uint32_t u = 0; u < 2; u++) {
                        const uint32_t bit = 1u << (w[u] & 31);
                        if (ok[u] && !(word[u] & bit) && (single || wc[u] == c[u]))
                            claimed[u] = !(atomicOr(&vis[w[u] >> 5], bit) & bit);
                    }
#pragma unroll
                    for (uint32_t u = 0; u < 2; u++) lpush(claimed[u], w[u], qout, cnt);
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            L.cnt[0] = L.cnt[1];
            L.cnt[2] = L.cnt[3];
        }
        __syncthreads();
        in ^= 1;
    }
}

// pivots: one per live class (first = single class, reduced in LDS)
template <int BS> __device__ uint32_t g_pivots(Lds &L, const G1 &g, bool first)
{
    if (first) {
        unsigned long long m = 0;
        for (uint32_t v = threadIdx.x + 1; v <= g.V; v += BS)
            if (g.cls[v] != DEAD) {
                const unsigned long long k = bkey(g.Fr, g.Br, v);
                m = k > m ? k : m;
            }
        for (int o = 32; o > 0; o >>= 1) {
            unsigned long long x = __shfl_xor(m, o);
            m = x > m ? x : m;
        }
        if (threadIdx.x == 0) L.key = 0;
        __syncthreads();
        if (ln() == 0 && m) atomicMax(&L.key, m);
        __syncthreads();
        if (threadIdx.x == 0) g.piv[0] = L.key;
    } else {
        // a wave whose live lanes share a class does one atomic
        unsigned long long *piv = reinterpret_cast<unsigned long long *>(g.piv);
        for (uint32_t base = threadIdx.x & ~63u; base <= g.V; base += BS) {
            const uint32_t v = base + ln();
            const uint32_t c = (v >= 1 && v <= g.V) ? g.cls[v] : DEAD;
            const bool ok = c != DEAD;
            const uint64_t live = __ballot(ok);
            if (!live) continue;
            const int lead = __ffsll((unsigned long long)live) - 1;
            const uint32_t c0 = __shfl(c, lead);
            unsigned long long k = ok ? bkey(g.Fr, g.Br, v) : 0ull;
            if (__all(!ok || c == c0)) {
                for (int o = 32; o > 0; o >>= 1) {
                    const unsigned long long x = __shfl_xor(k, o);
                    k = x > k ? x : k;
                }
                if ((int)ln() == lead && k > __hip_atomic_load(&piv[c0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    atomicMax(&piv[c0], k);
            } else if (ok && k > __hip_atomic_load(&piv[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                atomicMax(&piv[c], k);
            }
        }
    }
    if (threadIdx.x == 0) L.npiv = 0;
    __syncthreads();
    for (uint32_t v = threadIdx.x + 1; v <= g.V; v += BS) {
        const uint32_t c = g.cls[v];
        const bool win = c != DEAD && g.piv[c] == bkey(g.Fr, g.Br, v);
        if (win) {
            atomicOr(&g.fvis[v >> 5], 1u << (v & 31));
            atomicOr(&g.bvis[v >> 5], 1u << (v & 31));
        }
        lpush(win, v, g.pivlist, &L.npiv);
    }
    __syncthreads();
    const uint32_t n = L.npiv;
    for (uint32_t i = threadIdx.x; i < n; i += BS) {
        g.queue(0)[i] = g.pivlist[i];
        g.queue(2)[i] = g.pivlist[i];
    }
    if (threadIdx.x == 0) {
        L.cnt[0] = n;
        L.cnt[2] = n;
    }
    __syncthreads();
    return n;
}

// update (scc_kernels.cu:124-142) + clear the next round's pivot slots
template <int BS> __device__ uint32_t g_update(Lds &L, const G1 &g)
{
    uint32_t live = 0;
    for (uint32_t v = threadIdx.x + 1; v <= g.V; v += BS) {
        const uint32_t c = g.cls[v];
        if (c == DEAD) continue;
        uint32_t p = bkey_vertex(g.piv[c]);
        if (p - 1u >= g.V) {   // every live class elects a pivot; keep indices in range if not
            L.anom = 1;
            p = v;
        }
        const uint32_t f = bbit(g.fvis, v), b = bbit(g.bvis, v);
        if (f && b) {
            g.lab[v] = p;
            g.cls[v] = DEAD;
        } else {
            g.cls[v] = 3u * p + f + 2u * b;
            live++;
        }
    }
    const uint32_t n = wg_sum<BS>(L, live);
    for (uint32_t i = threadIdx.x; i < L.npiv; i += BS) {
        const uint32_t p = g.pivlist[i];
        g.piv[3ull * p] = 0;
        g.piv[3ull * p + 1] = 0;
        g.piv[3ull * p + 2] = 0;
    }
    __syncthreads();
    return n;
}

template <int BS> __device__ void g_pivot_round(Lds &L, const G1 &g, bool first, uint32_t W, uint32_t *live)
{
    g_init_vis<BS>(g);
    __syncthreads();
    g_pivots<BS>(L, g, first);
    g_reach<BS>(L, g, true, first, W);
    *live = g_update<BS>(L, g);
}

// "WCC" min reachable label with pointer jumping (scc_kernels.cu:477-536)
template <int BS> __device__ void g_minlabel(Lds &L, const G1 &g)
{
    for (uint32_t v = threadIdx.x + 1; v <= g.V; v += BS)
        if (g.cls[v] != DEAD) g.cls[v] = v;
    __syncthreads();
    for (uint32_t it = 0; it <= g.V; it++) {   // labels only decrease
        uint32_t changed = 0;
        for (uint32_t v = threadIdx.x + 1; v <= g.V; v += BS) {
            // label and row offsets together (an inactive lane folds an
            // empty range, so the offsets are used on every path)
            const uint32_t m0 = g.cls[v], b = g.Fr[v], e = g.Fr[v + 1];
            const bool act = m0 != DEAD;
            uint32_t m = g_fold<false>(g.Fc, b, act ? e : b, g.cls, m0);
            if (!act) continue;
            for (int it = 0; it < 32; it++) {
                const uint32_t x = g.cls[m];
                if (x >= m) break;
                m = x;
            }
            if (m < m0) {
                __hip_atomic_store(&g.cls[v], m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                changed = 1;
            }
        }
        if (!wg_sum<BS>(L, changed)) break;
    }
    for (uint32_t v = threadIdx.x + 1; v <= g.V; v += BS) {
        const uint32_t m = g.cls[v];
        if (m == DEAD) continue;
        if (m == v) g.piv[3ull * v] = 0;
        g.cls[v] = 3u * m;
    }
    __syncthreads();
}

// vSlota colouring loop (scc_Coloring.cu:590-637)
template <int BS> __device__ void g_color_loop(Lds &L, const G1 &g, uint32_t W, uint32_t live)
{
    for (uint32_t round = 0; live && round <= g.V; round++) {   // a round finishes >= 1 SCC
        for (uint32_t v = threadIdx.x; v <= g.V; v += BS) g.aux[v] = (v && g.cls[v] != DEAD) ? v : 0u;
        __syncthreads();
        for (uint32_t it = 0; it <= g.V; it++) {   // colours only increase
            uint32_t changed = 0;
            for (uint32_t v = threadIdx.x + 1; v <= g.V; v += BS) {
                // as in g_minlabel
                const uint32_t m0 = g.aux[v], b = g.Fr[v], e = g.Fr[v + 1];
                const bool act = m0 != 0;
                uint32_t m = g_fold<true>(g.Fc, b, act ? e : b, g.aux, m0);
                if (!act) continue;
                for (int it = 0; it < 32; it++) {
                    const uint32_t x = g.aux[m];
                    if (x <= m) break;
                    m = x;
                }
                if (m > m0) {
                    __hip_atomic_store(&g.aux[v], m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    changed = 1;
                }
            }
            if (!wg_sum<BS>(L, changed)) break;
        }
        g_init_vis<BS>(g);
        if (threadIdx.x == 0) L.cnt[0] = 0;
        __syncthreads();
        for (uint32_t v = threadIdx.x + 1; v <= g.V; v += BS) {
            bool root = false;
            if (g.cls[v] != DEAD) {
                const uint32_t c = g.aux[v];
                g.cls[v] = c;
                root = c == v;
                if (root) atomicOr(&g.fvis[v >> 5], 1u << (v & 31));
            }
            lpush(root, v, g.queue(0), &L.cnt[0]);
        }
        __syncthreads();
        g_reach<BS>(L, g, false, false, W);
        uint32_t lv = 0;
        for (uint32_t v = threadIdx.x + 1; v <= g.V; v += BS) {
            const uint32_t c = g.cls[v];
            if (c == DEAD) continue;
            if (bbit(g.fvis, v)) {
                g.lab[v] = c;
                g.cls[v] = DEAD;
            } else {
                lv++;
            }
        }
        live = wg_sum<BS>(L, lv);
    }
}

// canonical labels; returns the number of SCCs
// The minimum per label: lanes of a wave that share a label combine first
// (up to four labels per wave step, then lane by lane), and the guard reads
// the current minimum at agent scope (the CU's L1 may hold a stale copy),
// so a large SCC costs about one atomic per wave instead of one per vertex.
template <int BS> __device__ uint32_t g_canon(Lds &L, const G1 &g)
{
    for (uint32_t v = threadIdx.x; v <= g.V; v += BS) g.aux[v] = DEAD;
    __syncthreads();
    for (uint32_t base = threadIdx.x & ~63u; base <= g.V; base += BS) {
        const uint32_t v = base + ln();
        const uint32_t l = (v >= 1 && v <= g.V) ? g.lab[v] : LAB_FINAL;
        bool pending = !(l & LAB_FINAL);
        for (int it = 0; it < 4; it++) {
            const uint64_t pm = __ballot(pending);
            if (!pm) break;
            const int lead = __ffsll((unsigned long long)pm) - 1;
            const uint32_t rl = __shfl(l, lead);
            const bool mine = pending && l == rl;
            uint32_t m = mine ? v : DEAD;
            for (int o = 32; o > 0; o >>= 1) {
                const uint32_t x = __shfl_xor(m, o);
                m = x < m ? x : m;
            }
            if ((int)ln() == lead &&
                m < __hip_atomic_load(&g.aux[rl], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMin(&g.aux[rl], m);
            pending = pending && !mine;
        }
        if (pending && v < __hip_atomic_load(&g.aux[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMin(&g.aux[l], v);
    }
    __syncthreads();
    uint32_t roots = 0;
    for (uint32_t v = threadIdx.x + 1; v <= g.V; v += BS) {
        const uint32_t l = g.lab[v];
        const uint32_t r = (l & LAB_FINAL) ? (l & ~LAB_FINAL) : g.aux[l];
        g.lab[v] = r;
        roots += r == v;
    }
    if (threadIdx.x == 0) g.lab[0] = 0;
    return wg_sum<BS>(L, roots);
}

} // namespace

// LDS tiers.  A graph whose visited bitmaps fit `vis_words` words each keeps
// them in LDS, so probes and claims are LDS operations (a device-wide
// atomic costs a share of a chip-wide rate of about 10 G/s); a graph whose
// class words fit `cls_words` keeps those in LDS as well, so the
// neighbour-class tests of the trims, searches and colouring stay on the
// CU.  Larger graphs use their slots in global memory (the views are
// generic pointers, so the phase code is the same either way).  The tiers
// shrink with the workgroup so that smaller workgroups share a CU:
//   1,024 lanes: bitmaps up to 2^16 vertices, classes up to 2^15 (about
//                145 KB: one workgroup per CU, which its registers allow)
//     512 lanes: bitmaps up to 2^15, classes up to 2^14 (about 74 KB: two)
//     256 lanes: bitmaps up to 2^15, classes up to 2^13 (about 41 KB: three)
template <int BS> struct BTier;
template <> struct BTier<1024> {
    static constexpr uint32_t vis_words = 2 * ((65536 + 64) / 64);
    static constexpr uint32_t cls_words = (1u << 15) + 2;
};
template <> struct BTier<512> {
    static constexpr uint32_t vis_words = 2 * ((32768 + 64) / 64);
    static constexpr uint32_t cls_words = (1u << 14) + 2;
};
template <> struct BTier<256> {
    static constexpr uint32_t vis_words = 2 * ((32768 + 64) / 64);
    static constexpr uint32_t cls_words = (1u << 13) + 2;
};

template <int BS>
__global__ __launch_bounds__(BS) void batch_scc_kernel(BatchDev B, uint32_t first, uint32_t count)
{
    __shared__ Lds L;
    __shared__ uint32_t s_vis[2][BTier<BS>::vis_words];
    __shared__ uint32_t s_cls[BTier<BS>::cls_words];
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) L.graph = atomicAdd(B.next, 1u);
        __syncthreads();
        const uint32_t k = L.graph;
        if (k >= count) break;
        const uint32_t gi = B.order[first + k];
        const uint64_t vo = B.voff[gi], eo = B.eoff[gi];
        G1 g;
        g.V = B.Vn[gi];
        g.Fr = B.Fr + vo;
        g.Br = B.Br + vo;
        g.Fc = B.Fc + eo;
        g.Bc = B.Bc + eo;
        g.cls = g.V + 1 <= BTier<BS>::cls_words ? s_cls : B.cls + vo;
        g.lab = B.lab + vo;
        g.aux = B.aux + vo;
        // g_init_vis writes 2 * ((V + 64) / 64) words of each bitmap
        const bool lds_vis = 2 * ((g.V + 64) / 64) <= BTier<BS>::vis_words;
        g.fvis = lds_vis ? s_vis[0] : B.fvis + vo / 32;
        g.bvis = lds_vis ? s_vis[1] : B.bvis + vo / 32;
        g.piv = B.piv + 3 * vo;
        g.q = B.q + 4 * vo;
        g.pivlist = B.pivlist + vo;
        const uint32_t cfg = B.cfg[gi];
        const uint32_t W = cfg_lanes(cfg);
        const uint64_t t0 = wall_clock64();
        for (uint32_t v = threadIdx.x; v <= g.V; v += BS) {
            g.cls[v] = v ? 0u : DEAD;
            g.lab[v] = v ? (v | LAB_FINAL) : 0u;
        }
        if (threadIdx.x == 0) {
            g.piv[0] = 0;
            L.anom = 0;
        }
        __syncthreads();
        uint32_t live = g.V;
        if (!cfg_slota(cfg)) {
            if (cfg_t1(cfg)) g_trim1<BS>(L, g, true);
            g_pivot_round<BS>(L, g, true, W, &live);
            if (cfg_t1(cfg)) g_trim1<BS>(L, g, true);
            if (cfg_t2(cfg)) g_trim2<BS>(g);
            if (cfg_t1(cfg)) g_trim1<BS>(L, g, true);
            g_minlabel<BS>(L, g);
            for (uint32_t round = 0; round <= g.V + 1; round++) {
                g_pivot_round<BS>(L, g, false, W, &live);
                if (!live) break;
            }
        } else {
            if (cfg_t1(cfg)) g_trim1<BS>(L, g, false);
            g_pivot_round<BS>(L, g, true, W, &live);
            g_color_loop<BS>(L, g, W, live);
        }
        const uint32_t n = g_canon<BS>(L, g);
        if (threadIdx.x == 0) {
            B.nscc[gi] = L.anom ? BATCH_ANOMALY : n;
            B.tstart[gi] = t0;
            B.tend[gi] = wall_clock64();
        }
    }
}

uint32_t batch_cfg(int lanes, bool trim1, bool trim2, bool slota)
{
    uint32_t lg = 0;
    while ((1 << (lg + 1)) <= lanes && lg < 6) lg++;
    return lg | (trim1 ? 8u : 0u) | (trim2 ? 16u : 0u) | (slota ? 32u : 0u);
}

int batch_occupancy(int block)
{
    int n = 0;
    hipError_t e;
    if (block == 1024)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_scc_kernel<1024>, 1024, 0);
    else if (block == 512)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_scc_kernel<512>, 512, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_scc_kernel<256>, 256, 0);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return n < 1 ? 1 : n;
}

void k_batch(hipStream_t s, const BatchDev &B, uint32_t first, uint32_t count, int block, uint32_t grid)
{
    if (!count) return;
    if (block == 1024)
        batch_scc_kernel<1024><<<grid, 1024, 0, s>>>(B, first, count);
    else if (block == 512)
        batch_scc_kernel<512><<<grid, 512, 0, s>>>(B, first, count);
    else
        batch_scc_kernel<256><<<grid, 256, 0, s>>>(B, first, count);
}

} // namespace sccamd
