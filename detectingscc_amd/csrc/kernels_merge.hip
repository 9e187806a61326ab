// kernels_merge.hip -- partitioned local SCC and the cross-partition merge
// (detectSCC, scc_WCC.cu:1000-1617), rebuilt as a quotient-graph SCC.
//
// After the global trims, every live vertex starts in the class of its
// partition, so the ordinary per-class FW-BW rounds compute the SCCs of the
// union of partition-induced subgraphs ("local SCCs", scc_WCC.cu:1175-1269).
// Every local SCC lies inside one global SCC, so contracting each local SCC
// to a vertex and keeping every arc between different local SCCs gives a
// quotient graph whose SCCs are exactly the unions of local SCCs that form
// global SCCs.  This replaces the reference's host set-intersection merge
// (scc_WCC.cu:1276-1531, defect D7), which is not correct in general.
//
// Quotient vertex ids are assigned in increasing order of the local SCC's
// minimum vertex id, so the minimum quotient id of a quotient SCC maps back
// to the minimum vertex id of the global SCC: labels stay canonical.
#include "merge.h"

#include <algorithm>

namespace sccamd {

constexpr uint32_t SCAN_TILE = 1024;

// ---------------------------------------------------------------- scan ---
// Exclusive prefix sum of n u32 values (in-place allowed), totals < 2^32.
__global__ __launch_bounds__(256) void scan_tiles_kernel(const uint32_t *__restrict__ in, uint32_t *out,
                                                         uint32_t *tile_sums, uint64_t n)
{
    __shared__ uint32_t s_w[4];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint32_t x[4], t = 0;
    for (int k = 0; k < 4; k++) {
        const uint64_t i = base + threadIdx.x * 4 + k;
        x[k] = i < n ? in[i] : 0u;
        t += x[k];
    }
    // workgroup exclusive scan of t
    const uint32_t l = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t incl = t;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(incl, o);
        if (l >= (uint32_t)o) incl += y;
    }
    if (l == 63) s_w[w] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (uint32_t i = 0; i < 4; i++) {
        off += i < w ? s_w[i] : 0u;
        tot += s_w[i];
    }
    uint32_t run = off + incl - t;
    for (int k = 0; k < 4; k++) {
        const uint64_t i = base + threadIdx.x * 4 + k;
        if (i < n) out[i] = run;
        run += x[k];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

__global__ void scan_sums_kernel(uint32_t *tile_sums, uint32_t ntiles)
{
    // one workgroup; sequential over chunks of 256 with a wave scan each
    __shared__ uint32_t s_carry;
    __shared__ uint32_t s_w[4];
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < ntiles; b += 256) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t x = i < ntiles ? tile_sums[i] : 0u;
        const uint32_t l = threadIdx.x & 63u, w = threadIdx.x >> 6;
        uint32_t incl = x;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(incl, o);
            if (l >= (uint32_t)o) incl += y;
        }
        if (l == 63) s_w[w] = incl;
        __syncthreads();
        uint32_t off = 0, tot = 0;
        for (uint32_t k = 0; k < 4; k++) {
            off += k < w ? s_w[k] : 0u;
            tot += s_w[k];
        }
        const uint32_t carry = s_carry;
        if (i < ntiles) tile_sums[i] = carry + off + incl - x;
        __syncthreads();
        if (threadIdx.x == 0) s_carry = carry + tot;
        __syncthreads();
    }
}

__global__ void scan_add_kernel(uint32_t *out, const uint32_t *__restrict__ tile_sums, uint64_t n)
{
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    const uint32_t add = tile_sums[blockIdx.x];
    for (uint32_t k = threadIdx.x; k < SCAN_TILE; k += blockDim.x)
        if (base + k < n) out[base + k] += add;
}

void k_exscan(hipStream_t s, const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *tile_sums)
{
    const uint32_t ntiles = (uint32_t)((n + SCAN_TILE - 1) / SCAN_TILE);
    if (!ntiles) return;
    scan_tiles_kernel<<<ntiles, 256, 0, s>>>(in, out, tile_sums, n);
    scan_sums_kernel<<<1, 256, 0, s>>>(tile_sums, ntiles);
    scan_add_kernel<<<ntiles, 256, 0, s>>>(out, tile_sums, n);
}

uint64_t scan_tile_count(uint64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

// ------------------------------------------------------- local phase ---
// Live vertices enter the class of their (dense) partition id.
__global__ void local_init_kernel(uint32_t *cls, const uint32_t *__restrict__ part, uint64_t *piv,
                                  uint32_t *tbits, uint32_t V)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= V; v += gridDim.x * blockDim.x) {
        const bool dead = v == 0 || cls[v] == DEAD;
        if (!dead) {
            const uint32_t c = 3u * part[v];
            cls[v] = c;
            piv[c] = 0;
        }
        // remember who was finished before the local phase (global singletons)
        const uint64_t m = __ballot(dead);
        if ((threadIdx.x & 63u) == 0) {
            tbits[v >> 5] = (uint32_t)m;
            tbits[(v >> 5) + 1] = (uint32_t)(m >> 32);
        }
    }
}

void k_local_init(hipStream_t s, const DevGraph &g, DevState &st, uint32_t *tbits)
{
    // wave-aligned 1-D grid: each wave covers 64 consecutive ids
    const uint32_t blocks = (uint32_t)((((uint64_t)g.V + 1) + 255) / 256);
    local_init_kernel<<<blocks, 256, 0, s>>>(st.cls, st.part, st.piv, tbits, g.V);
}

__device__ __forceinline__ bool tbit(const uint32_t *tb, uint32_t v) { return (tb[v >> 5] >> (v & 31)) & 1u; }

// ---------------------------------------------------- quotient graph ---
// isrep[v] = 1 for the minimum vertex of each local SCC that took part in
// the local phase.
__global__ void q_mark_kernel(const uint32_t *__restrict__ lab, const uint32_t *__restrict__ tbits,
                              uint32_t *isrep, uint32_t V)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= V; v += gridDim.x * blockDim.x)
        isrep[v] = (v >= 1 && !tbit(tbits, v) && lab[v] == v) ? 1u : 0u;
}

// count (pass 0) or place (pass 1) every arc between different local SCCs
template <int PASS>
__global__ void q_arcs_kernel(const uint32_t *__restrict__ Fr, const uint32_t *__restrict__ Fc,
                              const uint32_t *__restrict__ lab, const uint32_t *__restrict__ tbits,
                              const uint32_t *__restrict__ qid, uint32_t *cnt_out, uint32_t *cnt_in,
                              uint32_t *Qc, uint32_t *QBc, uint32_t V)
{
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x + 1; u <= V; u += gridDim.x * blockDim.x) {
        if (tbit(tbits, u)) continue;
        const uint32_t ru = lab[u];
        const uint32_t qu = qid[ru] + 1;
        for (uint32_t i = Fr[u], e = Fr[u + 1]; i < e; i++) {
            const uint32_t w = Fc[i];
            if (tbit(tbits, w)) continue;
            const uint32_t rw = lab[w];
            if (rw == ru) continue;
            const uint32_t qw = qid[rw] + 1;
            if (PASS == 0) {
                atomicAdd(&cnt_out[qu], 1u);
                atomicAdd(&cnt_in[qw], 1u);
            } else {
                Qc[atomicAdd(&cnt_out[qu], 1u)] = qw;
                QBc[atomicAdd(&cnt_in[qw], 1u)] = qu;
            }
        }
    }
}

// Quotient id of every vertex: 1 + qid of its local SCC, 0 for a vertex
// that did not take part in the local phase (finished before it).
__global__ void q_vid_kernel(const uint32_t *__restrict__ lab, const uint32_t *__restrict__ tbits,
                             const uint32_t *__restrict__ qid, uint32_t *qv, uint32_t V)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= V; v += gridDim.x * blockDim.x) {
        // the finished bit and the label go out together; a finished
        // vertex's label (LAB_FINAL set) is not an index, vertex 0's stands in
        const bool fin = tbit(tbits, v);
        const uint32_t l = lab[v];
        const bool in = v >= 1 && !fin;
        const uint32_t q = qid[in ? l : 0u];
        qv[v] = in ? q + 1 : 0u;
    }
}

// Arcs between different local SCCs as 64-bit keys (source quotient id << 32
// | target quotient id), which two radix sorts turn into the quotient CSR.
// Counting into per-quotient-vertex rows took two atomics per arc on a few
// hundred addresses (a community graph's local SCCs) and serialised (3.3 ms
// per pass on 100M arcs).  Here a wave takes 64 consecutive vertices and
// their arcs, which are contiguous in Fc: lane l takes arcs A0 + l, A0 + l +
// 64, ... (QK_UNR of them in flight), finds the arc's source among the 64
// row starts with a 6-step search across lanes, and compares quotient ids.
// Counting (EMIT false) adds one atomic per workgroup; emitting appends the
// keys through a wave queue in LDS, one reservation per flush.
constexpr uint32_t QK_WQ = 512;   // keys per wave queue
constexpr uint32_t QK_UNR = 4;    // arcs per lane in flight

template <bool EMIT>
__global__ __launch_bounds__(256) void q_keys_kernel(const uint32_t *__restrict__ Fr, const uint32_t *__restrict__ Fc,
                                                     const uint32_t *__restrict__ qv, unsigned long long *n,
                                                     uint64_t *keys, uint64_t cap, uint32_t V)
{
    __shared__ uint64_t s_q[4][EMIT ? QK_WQ : 1];
    __shared__ unsigned long long s_sum[4];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    uint64_t *wq = s_q[wv];
    uint32_t nq = 0;
    unsigned long long cnt = 0;
    const uint32_t nwaves = gridDim.x * 4;
    for (uint32_t c = blockIdx.x * 4 + wv; (uint64_t)c * 64 <= V; c += nwaves) {
        const uint32_t base = c * 64, u = base + lane;
        const uint32_t rs = Fr[min(u, V + 1)];                      // row start (E past V)
        const uint32_t A1 = Fr[min(base + 64, V + 1)];
        const uint32_t qu_lane = u <= V ? qv[u] : 0u;
        const uint32_t A0 = __shfl(rs, 0);
        for (uint32_t j0 = A0; j0 < A1; j0 += 64 * QK_UNR) {
            uint32_t w[QK_UNR], qu[QK_UNR];
            bool ok[QK_UNR];
#pragma unroll
            for (uint32_t t = 0; t < QK_UNR; t++) {
                const uint32_t j = j0 + t * 64 + lane;
                ok[t] = j < A1;
                w[t] = Fc[ok[t] ? j : A1 - 1];   // A1 > A0: a valid arc
                // source: the last lane whose row start is <= j
                uint32_t lo = 0, hi = 63;
#pragma unroll
                for (int it = 0; it < 6; it++) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if ((uint32_t)__shfl(rs, (int)mid) <= j) lo = mid;
                    else hi = mid - 1;
                }
                qu[t] = __shfl(qu_lane, (int)lo);
            }
            uint32_t qw[QK_UNR];
#pragma unroll
            for (uint32_t t = 0; t < QK_UNR; t++) qw[t] = qv[w[t]];
            if (EMIT && nq + QK_UNR * 64 > QK_WQ) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                unsigned long long b0 = 0;
                if (lane == 0) b0 = atomicAdd(n, (unsigned long long)nq);
                b0 = __shfl(b0, 0);
                for (uint32_t i = lane; i < nq; i += 64)
                    if (b0 + i < cap) keys[b0 + i] = wq[i];
                nq = 0;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
#pragma unroll
            for (uint32_t t = 0; t < QK_UNR; t++) {
                const bool cross = ok[t] && qu[t] && qw[t] && qu[t] != qw[t];
                if (EMIT) {
                    const uint64_t m = __ballot(cross);
                    if (cross) wq[nq + (uint32_t)__popcll(m & ((1ull << lane) - 1))] = ((uint64_t)qu[t] << 32) | qw[t];
                    nq += (uint32_t)__popcll(m);
                } else {
                    cnt += cross;
                }
            }
        }
    }
    if (EMIT) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (nq) {
            unsigned long long b0 = 0;
            if (lane == 0) b0 = atomicAdd(n, (unsigned long long)nq);
            b0 = __shfl(b0, 0);
            for (uint32_t i = lane; i < nq; i += 64)
                if (b0 + i < cap) keys[b0 + i] = wq[i];
        }
    } else {
        for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
        if (lane == 0) s_sum[wv] = cnt;
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned long long t = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
            if (t) atomicAdd(n, t);
        }
    }
}

__global__ void q_rep_kernel(const uint32_t *__restrict__ isrep, const uint32_t *__restrict__ qid,
                             uint32_t *rep, uint32_t V)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x + 1; v <= V; v += gridDim.x * blockDim.x)
        if (isrep[v]) rep[qid[v]] = v;
}

// global label = minimum vertex of the quotient SCC's minimum local SCC
__global__ void q_map_kernel(uint32_t *lab, const uint32_t *__restrict__ tbits,
                             const uint32_t *__restrict__ qid, const uint32_t *__restrict__ qlab,
                             const uint32_t *__restrict__ rep, uint32_t *ctr, uint32_t V)
{
    uint32_t roots = 0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x + 1; v <= V; v += gridDim.x * blockDim.x) {
        uint32_t l = lab[v];
        if (!tbit(tbits, v) && qlab) l = rep[qlab[qid[l] + 1] - 1];
        lab[v] = l;
        roots += l == v;
    }
    for (int o = 32; o > 0; o >>= 1) roots += __shfl_xor(roots, o);
    // one add per wave, spread over the counter's slots
    if ((threadIdx.x & 63u) == 0 && roots)
        atomicAdd(&ctr[CSLOT(C_NSCC, (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % kSpread)], roots);
}

static uint32_t grid_for(uint64_t n)
{
    uint64_t b = (n + 255) / 256;
    return (uint32_t)(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

void k_q_mark(hipStream_t s, const DevGraph &g, const DevState &st, const uint32_t *tbits, uint32_t *isrep)
{
    q_mark_kernel<<<grid_for((uint64_t)g.V + 1), 256, 0, s>>>(st.lab, tbits, isrep, g.V);
}

void k_q_arcs(hipStream_t s, int pass, const DevGraph &g, const DevState &st, const uint32_t *tbits,
              const uint32_t *qid, uint32_t *cnt_out, uint32_t *cnt_in, uint32_t *Qc, uint32_t *QBc)
{
    if (pass == 0)
        q_arcs_kernel<0><<<grid_for(g.V), 256, 0, s>>>(g.Fr, g.Fc, st.lab, tbits, qid, cnt_out, cnt_in,
                                                       Qc, QBc, g.V);
    else
        q_arcs_kernel<1><<<grid_for(g.V), 256, 0, s>>>(g.Fr, g.Fc, st.lab, tbits, qid, cnt_out, cnt_in,
                                                       Qc, QBc, g.V);
}

void k_q_vid(hipStream_t s, const DevGraph &g, const DevState &st, const uint32_t *tbits, const uint32_t *qid,
             uint32_t *qv)
{
    q_vid_kernel<<<grid_for((uint64_t)g.V + 1), 256, 0, s>>>(st.lab, tbits, qid, qv, g.V);
}

void k_q_keys(hipStream_t s, bool emit, const DevGraph &g, const uint32_t *qv, unsigned long long *n,
              uint64_t *keys, uint64_t cap)
{
    const uint64_t chunks = (uint64_t)g.V / 64 + 1;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(2048, (chunks + 3) / 4);
    if (emit) q_keys_kernel<true><<<blocks, 256, 0, s>>>(g.Fr, g.Fc, qv, n, keys, cap, g.V);
    else q_keys_kernel<false><<<blocks, 256, 0, s>>>(g.Fr, g.Fc, qv, n, nullptr, 0, g.V);
}

void k_q_rep(hipStream_t s, const DevGraph &g, const uint32_t *isrep, const uint32_t *qid, uint32_t *rep)
{
    q_rep_kernel<<<grid_for(g.V), 256, 0, s>>>(isrep, qid, rep, g.V);
}

void k_q_map(hipStream_t s, const DevGraph &g, DevState &st, const uint32_t *tbits, const uint32_t *qid,
             const uint32_t *qlab, const uint32_t *rep)
{
    q_map_kernel<<<grid_for(g.V), 256, 0, s>>>(st.lab, tbits, qid, qlab, rep, st.ctr, g.V);
}

} // namespace sccamd
