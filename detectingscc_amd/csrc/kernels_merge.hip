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
