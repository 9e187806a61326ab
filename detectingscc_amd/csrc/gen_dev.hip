// gen_dev.hip -- graph generators on the device.  Every draw is the host
// generators' counter-based one (rng.h: the same rnd(seed, stream, index)
// values, thresholds and seeded permutation), so a graph generated here is
// bit-identical to scc_gen_* with the same arguments; the arcs are then
// sorted on the device into the reference CSR layout (rows by (source,
// target), in-lists by (target, source), duplicates kept, dummy vertex 0),
// and the graph becomes the context's uploaded graph without passing
// through host memory.
#include <algorithm>
#include <cstring>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "batch_engine.h"
#include "engine.h"
#include "generators.h"
#include "host_util.h"
#include "merge.h"
#include "rng.h"

namespace sccamd {

#define HIP_TRY(call)                                                                       \
    do {                                                                                    \
        hipError_t _e = (call);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail(SCC_ERR_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), \
                        __FILE__, __LINE__);                                                \
    } while (0)

constexpr uint32_t GEN_BLOCK = 256;

static uint32_t gen_blocks(uint64_t n)
{
    const uint64_t b = (n + GEN_BLOCK - 1) / GEN_BLOCK;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 8192));
}

// Appends (u, v) for every lane with `keep`, one atomic per wave; the order
// of a device arc list does not matter (the CSR build sorts it).
__device__ __forceinline__ void push_arc(bool keep, uint32_t u, uint32_t v, uint32_t *s, uint32_t *d,
                                         unsigned long long *n)
{
    const uint64_t m = __ballot(keep);
    if (!m) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    const uint32_t lane = threadIdx.x & 63u;
    unsigned long long base = 0;
    if ((int)lane == leader) base = atomicAdd(n, (unsigned long long)__popcll(m));
    base = __shfl(base, leader);
    if (keep) {
        const uint64_t at = base + (uint64_t)__popcll(m & ((1ull << lane) - 1));
        s[at] = u;
        d[at] = v;
    }
}

// ER draws (er_arcs): draw k is (1 + below(rnd(seed, 1, 2k), V),
// 1 + below(rnd(seed, 1, 2k + 1), V)); self-loops are dropped.
__global__ void er_kernel(uint32_t V, uint64_t draws, uint64_t seed, uint32_t *s, uint32_t *d,
                          unsigned long long *n)
{
    for (uint64_t k0 = (uint64_t)blockIdx.x * blockDim.x; k0 < draws; k0 += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = k0 + threadIdx.x;
        uint32_t u = 0, v = 0;
        if (k < draws) {
            u = 1 + (uint32_t)below(rnd(seed, 1, 2 * k), V);
            v = 1 + (uint32_t)below(rnd(seed, 1, 2 * k + 1), V);
        }
        push_arc(k < draws && u != v, u, v, s, d, n);
    }
}

// Planted cycles (planted_cycles): cycle j has length
// L_j = cmin + below(rnd(seed, 2, j), cmax - cmin + 1) and visits the first
// L_j distinct values of 1 + below(rnd(seed, 3 + j, t), V), t = 0, 1, ...
__global__ void cycle_len_kernel(uint32_t cycles, uint32_t cmin, uint32_t cmax, uint64_t seed, uint32_t *len)
{
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < cycles; j += gridDim.x * blockDim.x)
        len[j] = cmin + (uint32_t)below(rnd(seed, 2, j), (uint64_t)(cmax - cmin + 1));
}

// One thread per cycle; the cycle's vertices are collected in its slice of
// `s` (a linear scan rejects repeats, as the host's set does), then its
// arcs written in place.
__global__ void cycle_fill_kernel(uint32_t V, uint32_t cycles, uint64_t seed, const uint32_t *len,
                                  const uint64_t *off, uint32_t *s, uint32_t *d)
{
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < cycles; j += gridDim.x * blockDim.x) {
        const uint32_t L = len[j];
        uint32_t *cyc = s + off[j];
        uint32_t have = 0;
        for (uint64_t t = 0; have < L; t++) {
            const uint32_t x = 1 + (uint32_t)below(rnd(seed, 3 + (uint64_t)j, t), V);
            bool seen = false;
            for (uint32_t i = 0; i < have && !seen; i++) seen = cyc[i] == x;
            if (!seen) cyc[have++] = x;
        }
        for (uint32_t i = 0; i < L; i++) d[off[j] + i] = cyc[(i + 1) % L];
    }
}

// R-MAT (rmat_arcs): arc i from rmat_arc, both endpoints permuted.
__global__ void rmat_kernel(RmatParams P, uint64_t seed, SeededPerm perm, int permute, uint32_t *s, uint32_t *d)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P.E; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t r, q;
        rmat_arc(P, seed, i, r, q);
        if (permute) {
            r = perm_at(perm, r);
            q = perm_at(perm, q);
        }
        s[i] = r + 1;
        d[i] = q + 1;
    }
}

// Grid (grid_arcs): vertex x + 1 at row x / side, column x % side keeps
// each of its N, W, E, S arcs inside the grid with probability p, drawn
// from rnd(seed, 20, 4 (x + 1) + k).
__global__ void grid_kernel(uint32_t side, double p, uint64_t seed, uint32_t *s, uint32_t *d,
                            unsigned long long *n)
{
    const uint64_t V = (uint64_t)side * side;
    for (uint64_t x0 = (uint64_t)blockIdx.x * blockDim.x; x0 < V; x0 += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t x = x0 + threadIdx.x;
        const bool in = x < V;
        const uint64_t r = in ? x / side : 0, c = in ? x % side : 0, id = x + 1;
        const bool nb = in && r > 0 && bernoulli(rnd(seed, 20, id * 4 + 0), p);
        const bool wb = in && c > 0 && bernoulli(rnd(seed, 20, id * 4 + 1), p);
        const bool eb = in && c + 1 < side && bernoulli(rnd(seed, 20, id * 4 + 2), p);
        const bool sb = in && r + 1 < side && bernoulli(rnd(seed, 20, id * 4 + 3), p);
        const uint32_t u = (uint32_t)id;
        push_arc(nb, u, u - side, s, d, n);
        push_arc(wb, u, u - 1, s, d, n);
        push_arc(eb, u, u + 1, s, d, n);
        push_arc(sb, u, u + side, s, d, n);
    }
}

// Planted communities (scc_gen_communities): community c holds the
// positions [c V / K, (c + 1) V / K) of the seeded permutation (position ->
// vertex - 1); part[v + 1] = the community of vertex v + 1.
__device__ __forceinline__ uint32_t comm_lo(uint64_t c, uint32_t V, uint32_t K) { return (uint32_t)(c * V / K); }

__global__ void community_part_kernel(uint32_t V, uint32_t K, SeededPerm perm, uint32_t *part)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const uint32_t pos = perm_inv(perm, v);
        uint32_t c = (uint32_t)(((uint64_t)pos * K) / V);
        while (c + 1 < K && pos >= comm_lo(c + 1, V, K)) c++;
        while (c > 0 && pos < comm_lo(c, V, K)) c--;
        part[v + 1] = c;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        part[0] = 0;
        part[V + 1] = 0;
    }
}

__global__ void community_arcs_kernel(uint32_t V, uint32_t K, double inter, uint64_t seed, SeededPerm perm,
                                      const uint32_t *part, uint64_t E, uint32_t *s, uint32_t *d)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = (uint32_t)below(rnd(seed, 30, i), V);
        uint32_t v;
        if (bernoulli(rnd(seed, 31, i), inter)) {
            v = (uint32_t)below(rnd(seed, 32, i), V);
        } else {
            const uint32_t c = part[u + 1];
            const uint32_t a0 = comm_lo(c, V, K), n = comm_lo((uint64_t)c + 1, V, K) - a0;
            v = perm_at(perm, a0 + (uint32_t)below(rnd(seed, 32, i), n));
        }
        s[i] = u + 1;
        d[i] = v + 1;
    }
}

// ---------------------------------------------------------- CSR build ---

__global__ void pack_keys_kernel(const uint32_t *hi, const uint32_t *lo, uint64_t E, uint64_t *keys)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E; i += (uint64_t)gridDim.x * blockDim.x)
        keys[i] = ((uint64_t)hi[i] << 32) | lo[i];
}

// From keys sorted ascending: C[i] = low word, and R[u] = the number of
// keys whose high word is below u, for u = 0..V+1 (position i fills the
// rows between the previous key's high word and its own).
__global__ void rows_from_keys_kernel(const uint64_t *keys, uint64_t E, uint32_t V, uint32_t *R, uint32_t *C)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= E; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t lo_u = i == 0 ? 0 : (keys[i - 1] >> 32) + 1;
        const uint64_t hi_u = i == E ? (uint64_t)V + 1 : (keys[i] >> 32);
        for (uint64_t u = lo_u; u <= hi_u; u++) R[u] = (uint32_t)i;
        if (i < E) C[i] = (uint32_t)keys[i];
    }
}

// Bits of a vertex id of a graph of V vertices (ids 1..V).
static uint32_t id_bits(uint32_t V)
{
    uint32_t bits = 1;
    while (bits < 32 && (1ull << bits) <= V) bits++;
    return bits;
}

// Sorted (hi, lo) keys of E arcs into the CSR (R, C) of V vertices.
static scc_status sort_rows(hipStream_t st, const uint32_t *hi, const uint32_t *lo, uint64_t E, uint32_t V,
                            uint64_t *kin, uint64_t *kout, void *tmp, size_t tmp_bytes, uint32_t *R, uint32_t *C)
{
    const uint32_t bits = id_bits(V);
    if (E) {
        pack_keys_kernel<<<gen_blocks(E), GEN_BLOCK, 0, st>>>(hi, lo, E, kin);
        size_t need = tmp_bytes;
        HIP_TRY(rocprim::radix_sort_keys(tmp, need, kin, kout, (unsigned int)E, 0u, 32u + bits, st));
    }
    rows_from_keys_kernel<<<gen_blocks(E + 1), GEN_BLOCK, 0, st>>>(kout, E, V, R, C);
    return SCC_OK;
}

// From keys sorted ascending: R[u] = the number of keys whose high word is
// below u (a binary search per u, so a graph with far more vertices than
// arcs costs no long serial fill), and C[i] = key i's low word.
__global__ void rows_by_search_kernel(const uint64_t *keys, uint64_t n, uint32_t V, uint32_t *R, uint32_t *C)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u <= (uint64_t)V + 1; u += stride) {
        uint64_t lo = 0, hi = n;   // first key with high word >= u
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if ((keys[mid] >> 32) < u) lo = mid + 1;
            else hi = mid;
        }
        R[u] = (uint32_t)lo;
    }
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) C[i] = (uint32_t)keys[i];
}

__global__ void swap_keys_kernel(const uint64_t *in, uint64_t n, uint64_t *out)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = (in[i] << 32) | (in[i] >> 32);
}

size_t csr_sort_bytes(uint32_t V, uint64_t n)
{
    size_t b = 0;
    (void)rocprim::radix_sort_keys(nullptr, b, (uint64_t *)nullptr, (uint64_t *)nullptr,
                                   (unsigned int)std::max<uint64_t>(n, 1), 0u, 32u + id_bits(V), (hipStream_t)0);
    return std::max<size_t>(b, 1);
}

scc_status csr_from_keys(hipStream_t st, uint32_t V, uint64_t n, uint64_t *keys, uint64_t *out, void *tmp,
                         size_t tmp_bytes, uint32_t *R, uint32_t *C, uint32_t *RB, uint32_t *CB)
{
    if (n > 0xFFFFFFFFull) return fail(SCC_ERR_INVALID, "too many arcs (%llu)", (unsigned long long)n);
    const uint32_t bits = id_bits(V);
    size_t need = tmp_bytes;
    if (n) HIP_TRY(rocprim::radix_sort_keys(tmp, need, keys, out, (unsigned int)n, 0u, 32u + bits, st));
    rows_by_search_kernel<<<gen_blocks(std::max<uint64_t>(n, (uint64_t)V + 2)), GEN_BLOCK, 0, st>>>(out, n, V, R, C);
    if (n) {
        // reverse keys from the sorted ones (the sort's input is not reused)
        swap_keys_kernel<<<gen_blocks(n), GEN_BLOCK, 0, st>>>(out, n, keys);
        need = tmp_bytes;
        HIP_TRY(rocprim::radix_sort_keys(tmp, need, keys, out, (unsigned int)n, 0u, 32u + bits, st));
    }
    rows_by_search_kernel<<<gen_blocks(std::max<uint64_t>(n, (uint64_t)V + 2)), GEN_BLOCK, 0, st>>>(out, n, V, RB, CB);
    return SCC_OK;
}

// A device arc list (1-based ids) into the engine's graph.
static scc_status adopt_arcs(Engine &eng, uint32_t V, uint64_t E, const uint32_t *s, const uint32_t *d)
{
    if (E > 0xFFFFFFFFull) return fail(SCC_ERR_INVALID, "too many arcs (%llu)", (unsigned long long)E);
    uint64_t *keys = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    HIP_TRY(rocprim::radix_sort_keys(nullptr, tmp_bytes, (uint64_t *)nullptr, (uint64_t *)nullptr,
                                     (unsigned int)std::max<uint64_t>(E, 1), 0u, 32u + id_bits(V), eng.stream));
    if (hipMalloc(&keys, sizeof(uint64_t) * 2 * std::max<uint64_t>(E, 1)) != hipSuccess ||
        hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)) != hipSuccess) {
        (void)hipGetLastError();
        if (keys) (void)hipFree(keys);
        return fail(SCC_ERR_NOMEM, "Error: Not enough memory on GPU to sort %llu arcs", (unsigned long long)E);
    }
    scc_status st = eng.adopt_begin(V, (uint32_t)E);
    if (st == SCC_OK)
        st = sort_rows(eng.stream, s, d, E, V, keys, keys + E, tmp, tmp_bytes, eng.g.Fr, eng.g.Fc);
    if (st == SCC_OK)
        st = sort_rows(eng.stream, d, s, E, V, keys, keys + E, tmp, tmp_bytes, eng.g.Br, eng.g.Bc);
    if (st == SCC_OK && hipStreamSynchronize(eng.stream) != hipSuccess)
        st = fail(SCC_ERR_DEVICE, "CSR build on the device failed: %s", hipGetErrorString(hipGetLastError()));
    (void)hipFree(keys);
    (void)hipFree(tmp);
    if (st == SCC_OK) eng.adopt_end();
    return st;
}

// Device arc buffers of n arcs (and an arc counter), freed on scope exit.
struct ArcBuf {
    uint32_t *s = nullptr, *d = nullptr;
    unsigned long long *n = nullptr;
    ~ArcBuf()
    {
        if (s) (void)hipFree(s);
        if (d) (void)hipFree(d);
        if (n) (void)hipFree(n);
    }
    scc_status alloc(uint64_t cap, hipStream_t st)
    {
        if (hipMalloc(&s, sizeof(uint32_t) * std::max<uint64_t>(cap, 1)) != hipSuccess ||
            hipMalloc(&d, sizeof(uint32_t) * std::max<uint64_t>(cap, 1)) != hipSuccess ||
            hipMalloc(&n, sizeof(unsigned long long)) != hipSuccess) {
            (void)hipGetLastError();
            return fail(SCC_ERR_NOMEM, "Error: Not enough memory on GPU for %llu arcs", (unsigned long long)cap);
        }
        HIP_TRY(hipMemsetAsync(n, 0, sizeof(unsigned long long), st));
        return SCC_OK;
    }
    scc_status count(hipStream_t st, uint64_t *out)
    {
        unsigned long long h = 0;
        HIP_TRY(hipMemcpyAsync(&h, n, sizeof h, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        *out = h;
        return SCC_OK;
    }
};

// ------------------------------------------------------ batch mix ---
// The 4,096-graph mix on the device.  Every graph's arcs become 64-bit keys
// (graph << 2b | source << b | target, b bits per vertex id) in one array
// (and the reverse keys in another); ER draws that are self-loops become
// the largest key, which sorts after every real one.  Two radix sorts then
// give every graph's columns in the reference order, and binary searches
// over the sorted keys give the graphs' column offsets and row arrays.
struct MixJob {
    uint64_t start;     // arc jobs: first item in the dense numbering of all arc-job items
    uint64_t pos;       // position of the job's first item in the key arrays
    uint64_t count;     // items
    uint64_t seed;
    uint32_t g, V;      // graph and its vertex count
    uint32_t kind;      // 0 ER draws, 1 R-MAT arcs, 2 one planted cycle
    uint32_t extra;     // cycle: its index j
    RmatParams P;       // R-MAT only
    SeededPerm perm;    // R-MAT only
};

__device__ __forceinline__ uint64_t mix_key(uint64_t g, uint32_t u, uint32_t v, uint32_t b)
{
    return (g << (2 * b)) | ((uint64_t)u << b) | v;
}

// Job of item x: the last job whose start is <= x (jobs sorted by start).
__device__ __forceinline__ uint32_t job_of(const MixJob *jobs, uint32_t njobs, uint64_t x)
{
    uint32_t lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (jobs[mid].start <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ER draws and R-MAT arcs: one item per thread.
__global__ void mix_arcs_kernel(const MixJob *jobs, uint32_t njobs, uint64_t total, uint32_t b, uint64_t *fwd,
                                uint64_t *rev)
{
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total;
         x += (uint64_t)gridDim.x * blockDim.x) {
        const MixJob &J = jobs[job_of(jobs, njobs, x)];
        const uint64_t k = x - J.start, at = J.pos + k;
        uint32_t u, v;
        if (J.kind == 0) {
            u = 1 + (uint32_t)below(rnd(J.seed, 1, 2 * k), J.V);
            v = 1 + (uint32_t)below(rnd(J.seed, 1, 2 * k + 1), J.V);
        } else {
            rmat_arc(J.P, J.seed, k, u, v);
            u = perm_at(J.perm, u) + 1;
            v = perm_at(J.perm, v) + 1;
        }
        const bool keep = u != v || J.kind != 0;   // ER drops self-loops; R-MAT keeps them
        fwd[at] = keep ? mix_key(J.g, u, v, b) : ~0ull;
        rev[at] = keep ? mix_key(J.g, v, u, b) : ~0ull;
    }
}

// Planted cycles: one thread per cycle; its vertices are collected in
// `tmp` at the cycle's items, a linear scan rejecting repeats.
__global__ void mix_cycles_kernel(const MixJob *jobs, uint32_t njobs, uint32_t b, uint32_t *tmp, uint64_t *fwd,
                                  uint64_t *rev)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < njobs; i += gridDim.x * blockDim.x) {
        const MixJob &J = jobs[i];
        const uint32_t L = (uint32_t)J.count;
        uint32_t *cyc = tmp + J.pos;
        uint32_t have = 0;
        for (uint64_t t = 0; have < L; t++) {
            const uint32_t x = 1 + (uint32_t)below(rnd(J.seed, 3 + (uint64_t)J.extra, t), J.V);
            bool seen = false;
            for (uint32_t q = 0; q < have && !seen; q++) seen = cyc[q] == x;
            if (!seen) cyc[have++] = x;
        }
        for (uint32_t q = 0; q < L; q++) {
            const uint32_t u = cyc[q], v = cyc[(q + 1) % L];
            fwd[J.pos + q] = mix_key(J.g, u, v, b);
            rev[J.pos + q] = mix_key(J.g, v, u, b);
        }
    }
}

__device__ __forceinline__ uint64_t lower_bound_key(const uint64_t *keys, uint64_t lo, uint64_t hi, uint64_t key)
{
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (keys[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// start[g] = the first sorted key of graph g (g = 0..G; start[G] = the
// number of real arcs).
__global__ void mix_starts_kernel(const uint64_t *keys, uint64_t n, uint32_t G, uint32_t b, uint64_t *start)
{
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g <= G; g += gridDim.x * blockDim.x)
        start[g] = lower_bound_key(keys, 0, n, (uint64_t)g << (2 * b));
}

// Row arrays of every graph (V_g + 2 entries at roff[g], offsets local to
// the graph's columns) and the columns themselves.
__global__ void mix_rows_kernel(const uint64_t *keys, const uint64_t *start, const uint64_t *roff, uint32_t G,
                                uint32_t b, uint32_t *R)
{
    const uint64_t total = roff[G];
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total;
         x += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = G - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (roff[mid] <= x) lo = mid;
            else hi = mid - 1;
        }
        const uint32_t g = lo, u = (uint32_t)(x - roff[g]);
        R[x] = (uint32_t)(lower_bound_key(keys, start[g], start[g + 1], mix_key(g, u, 0, b)) - start[g]);
    }
}

__global__ void mix_cols_kernel(const uint64_t *keys, uint64_t n, uint32_t b, uint32_t *C)
{
    const uint64_t mask = (1ull << b) - 1;
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (uint64_t)gridDim.x * blockDim.x)
        C[x] = (uint32_t)(keys[x] & mask);
}

// Device buffers of one generation, freed on scope exit.
struct DevBufs {
    std::vector<void *> p;
    ~DevBufs()
    {
        for (void *x : p)
            if (x) (void)hipFree(x);
    }
    template <class T> scc_status get(T **out, size_t n)
    {
        void *x = nullptr;
        if (hipMalloc(&x, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) {
            (void)hipGetLastError();
            return fail(SCC_ERR_NOMEM, "Error: Not enough memory on GPU for the batch generation");
        }
        p.push_back(x);
        *out = static_cast<T *>(x);
        return SCC_OK;
    }
};

} // namespace sccamd

using namespace sccamd;

#define CTX_DEV(ctx)                                                                        \
    do {                                                                                    \
        clear_error();                                                                      \
        if (!(ctx)) return fail(SCC_ERR_INVALID, "null context");                           \
        HIP_TRY(hipSetDevice((ctx)->eng.device));                                           \
    } while (0)

extern "C" {

scc_status scc_gen_er_dev(scc_ctx *ctx, uint32_t V, uint64_t draws, uint32_t cycles, uint32_t cmin, uint32_t cmax,
                          uint64_t seed)
{
    CTX_DEV(ctx);
    if (V < 1 || V >= (1u << 30)) return fail(SCC_ERR_INVALID, "bad ER arguments");
    Engine &e = ctx->eng;
    hipStream_t st = e.stream;
    // planted_cycles' clamping
    if (cmax > V) cmax = V;
    if (cmin < 2) cmin = 2;
    if (cmin > cmax) cycles = 0;
    std::vector<uint32_t> len(cycles);
    std::vector<uint64_t> off((size_t)cycles + 1, 0);
    uint32_t *dlen = nullptr;
    uint64_t *doff = nullptr;
    if (cycles) {
        HIP_TRY(hipMalloc(&dlen, sizeof(uint32_t) * cycles));
        cycle_len_kernel<<<gen_blocks(cycles), GEN_BLOCK, 0, st>>>(cycles, cmin, cmax, seed, dlen);
        HIP_TRY(hipMemcpyAsync(len.data(), dlen, sizeof(uint32_t) * cycles, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (uint32_t j = 0; j < cycles; j++) off[j + 1] = off[j] + len[j];
    }
    ArcBuf a;
    scc_status s = a.alloc(draws + off[cycles], st);
    if (s == SCC_OK) {
        er_kernel<<<gen_blocks(draws), GEN_BLOCK, 0, st>>>(V, draws, seed, a.s, a.d, a.n);
        uint64_t n = 0;
        s = a.count(st, &n);
        if (s == SCC_OK && cycles) {
            if (hipMalloc(&doff, sizeof(uint64_t) * cycles) != hipSuccess) {
                (void)hipGetLastError();
                s = fail(SCC_ERR_NOMEM, "device allocation failed");
            } else {
                for (uint32_t j = 0; j < cycles; j++) off[j] += n;   // after the ER arcs
                if (hipMemcpyAsync(doff, off.data(), sizeof(uint64_t) * cycles, hipMemcpyHostToDevice, st) !=
                    hipSuccess)
                    s = fail(SCC_ERR_DEVICE, "copy failed");
                else
                    cycle_fill_kernel<<<gen_blocks(cycles), 64, 0, st>>>(V, cycles, seed, dlen, doff, a.s, a.d);
                n = off[cycles - 1] + len[cycles - 1];
            }
        }
        if (s == SCC_OK) s = adopt_arcs(e, V, n, a.s, a.d);
    }
    if (dlen) (void)hipFree(dlen);
    if (doff) (void)hipFree(doff);
    return s;
}

scc_status scc_gen_rmat_dev(scc_ctx *ctx, uint32_t scale, uint64_t E, double a, double b, double c, uint64_t seed)
{
    CTX_DEV(ctx);
    if (scale < 1 || scale > 29 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return fail(SCC_ERR_INVALID, "bad R-MAT arguments");
    const uint32_t V = 1u << scale;
    ArcBuf buf;
    scc_status s = buf.alloc(E, ctx->eng.stream);
    if (s != SCC_OK) return s;
    RmatParams P;
    P.ta = (uint32_t)std::lround(a * 65536.0);
    P.tab = (uint32_t)std::lround((a + b) * 65536.0);
    P.tabc = (uint32_t)std::lround((a + b + c) * 65536.0);
    P.scale = scale;
    P.words = (scale + 3) / 4;
    P.vcap = V;
    P.E = E;
    rmat_kernel<<<gen_blocks(E), GEN_BLOCK, 0, ctx->eng.stream>>>(P, seed, make_perm(V, seed), 1, buf.s, buf.d);
    return adopt_arcs(ctx->eng, V, E, buf.s, buf.d);
}

scc_status scc_gen_grid_dev(scc_ctx *ctx, uint32_t side, double p, uint64_t seed)
{
    CTX_DEV(ctx);
    if (side < 1 || (uint64_t)side * side >= (1u << 30) || p < 0 || p > 1)
        return fail(SCC_ERR_INVALID, "bad grid arguments");
    const uint64_t V = (uint64_t)side * side;
    ArcBuf a;
    scc_status s = a.alloc(4 * V, ctx->eng.stream);
    if (s != SCC_OK) return s;
    grid_kernel<<<gen_blocks(V), GEN_BLOCK, 0, ctx->eng.stream>>>(side, p, seed, a.s, a.d, a.n);
    uint64_t n = 0;
    if ((s = a.count(ctx->eng.stream, &n)) != SCC_OK) return s;
    return adopt_arcs(ctx->eng, (uint32_t)V, n, a.s, a.d);
}

scc_status scc_gen_communities_dev(scc_ctx *ctx, uint32_t V, uint32_t K, double deg, double inter, uint64_t seed,
                                   uint32_t *part)
{
    CTX_DEV(ctx);
    if (V < 1 || K < 1 || K > V || V >= (1u << 30) || deg < 0 || inter < 0 || inter > 1)
        return fail(SCC_ERR_INVALID, "bad community arguments");
    const uint64_t E = (uint64_t)std::llround(deg * V);
    hipStream_t st = ctx->eng.stream;
    ArcBuf a;
    scc_status s = a.alloc(E, st);
    if (s != SCC_OK) return s;
    uint32_t *dpart = nullptr;
    HIP_TRY(hipMalloc(&dpart, sizeof(uint32_t) * ((size_t)V + 2)));
    const SeededPerm perm = make_perm(V, seed);
    community_part_kernel<<<gen_blocks(V), GEN_BLOCK, 0, st>>>(V, K, perm, dpart);
    community_arcs_kernel<<<gen_blocks(E), GEN_BLOCK, 0, st>>>(V, K, inter, seed, perm, dpart, E, a.s, a.d);
    std::vector<uint32_t> hpart((size_t)V + 2);
    if (hipMemcpyAsync(hpart.data(), dpart, sizeof(uint32_t) * hpart.size(), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        s = fail(SCC_ERR_DEVICE, "community generation failed");
    (void)hipFree(dpart);
    if (s == SCC_OK) s = adopt_arcs(ctx->eng, V, E, a.s, a.d);
    if (s == SCC_OK) s = ctx->eng.upload_partition(hpart.data());
    if (s == SCC_OK && part) std::memcpy(part, hpart.data(), sizeof(uint32_t) * hpart.size());
    return s;
}

scc_status scc_gen_batch_mix_dev(scc_ctx *ctx, uint32_t num_graphs, uint32_t lgmin, uint32_t lgmax, double dmin,
                                 double dmax, uint64_t seed, scc_batch_host *copy)
{
    CTX_DEV(ctx);
    if (!num_graphs || lgmin < 1 || lgmax < lgmin || lgmax > 24 || dmin < 0 || dmax < dmin)
        return fail(SCC_ERR_INVALID, "bad batch-mix arguments");
    const uint32_t G = num_graphs;
    hipStream_t st = ctx->eng.stream;
    // per-graph draws on the host (they use pow and rounding, which the
    // device need not reproduce bit for bit); jobs for the device
    std::vector<uint32_t> V(G);
    std::vector<MixJob> arcs, cycles;
    uint64_t n = 0;
    uint32_t vmax = 2;
    for (uint32_t i = 0; i < G; i++) {
        const MixParams m = mix_params(seed, i, lgmin, lgmax, dmin, dmax);
        V[i] = m.V;
        vmax = std::max(vmax, m.V);
        MixJob J;
        std::memset(&J, 0, sizeof J);
        J.g = i;
        J.V = m.V;
        if (m.kind == 0 || m.kind == 1) {
            J.kind = m.kind;
            J.pos = n;
            J.count = m.E;
            J.seed = m.gs;
            if (m.kind == 1) {
                uint32_t scale = 1;
                while ((1ull << scale) < m.V) scale++;
                J.P = rmat_params(scale, m.V, m.E, 0.57, 0.19, 0.19);
                J.perm = make_perm(m.V, m.gs);
            }
            if (J.count) arcs.push_back(J);
            n += m.E;
        } else {
            uint64_t have = 0;
            const uint32_t nc = mix_cycles(m);
            for (uint32_t j = 0; j < nc; j++) {
                MixJob C = J;
                C.kind = 2;
                C.pos = n;
                C.count = mix_cycle_len(m, j);
                C.seed = m.gs;
                C.extra = j;
                cycles.push_back(C);
                n += C.count;
                have += C.count;
            }
            if (have < m.E) {
                J.kind = 0;
                J.pos = n;
                J.count = m.E - have;
                J.seed = m.gs ^ kMixFillSalt;
                arcs.push_back(J);
                n += J.count;
            }
        }
    }
    uint32_t b = 1;
    while ((1ull << b) <= vmax) b++;
    uint32_t gbits = 1;
    while ((1ull << gbits) <= G) gbits++;
    if (2 * b + gbits > 63) return fail(SCC_ERR_INVALID, "batch too large for 64-bit arc keys");
    if (n > 0xFFFFFFFFull) return fail(SCC_ERR_INVALID, "batch has too many arcs");
    std::vector<uint64_t> roff((size_t)G + 1, 0);
    for (uint32_t i = 0; i < G; i++) roff[i + 1] = roff[i] + V[i] + 2;
    DevBufs B;
    uint64_t *fwd, *rev, *fout, *rout, *dstart, *droff;
    uint32_t *tmp, *Fr, *Br, *Fc, *Bc;
    MixJob *djobs, *dcyc;
    scc_status s;
    const size_t N = std::max<uint64_t>(n, 1);
    if ((s = B.get(&fwd, N)) || (s = B.get(&rev, N)) || (s = B.get(&fout, N)) || (s = B.get(&rout, N)) ||
        (s = B.get(&tmp, N)) || (s = B.get(&dstart, (size_t)G + 1)) || (s = B.get(&droff, (size_t)G + 1)) ||
        (s = B.get(&Fr, roff[G])) || (s = B.get(&Br, roff[G])) || (s = B.get(&djobs, arcs.size())) ||
        (s = B.get(&dcyc, cycles.size())))
        return s;
    // the arc jobs' items numbered densely in job order
    uint64_t arc_items = 0;
    for (auto &J : arcs) {
        J.start = arc_items;
        arc_items += J.count;
    }
    if (!arcs.empty())
        HIP_TRY(hipMemcpyAsync(djobs, arcs.data(), sizeof(MixJob) * arcs.size(), hipMemcpyHostToDevice, st));
    if (!cycles.empty())
        HIP_TRY(hipMemcpyAsync(dcyc, cycles.data(), sizeof(MixJob) * cycles.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(droff, roff.data(), sizeof(uint64_t) * (G + 1), hipMemcpyHostToDevice, st));
    if (arc_items)
        mix_arcs_kernel<<<gen_blocks(arc_items), GEN_BLOCK, 0, st>>>(djobs, (uint32_t)arcs.size(), arc_items, b,
                                                                    fwd, rev);
    if (!cycles.empty())
        mix_cycles_kernel<<<gen_blocks(cycles.size()), 64, 0, st>>>(dcyc, (uint32_t)cycles.size(), b, tmp, fwd,
                                                                    rev);
    // sort both key arrays
    const uint32_t bits = 2 * b + gbits;
    size_t tmp_bytes = 0;
    HIP_TRY(rocprim::radix_sort_keys(nullptr, tmp_bytes, fwd, fout, (unsigned int)N, 0u, bits, st));
    void *sort_tmp;
    if ((s = B.get((uint8_t **)&sort_tmp, tmp_bytes))) return s;
    if (n) {
        // valid keys are below 2^bits; a dropped draw's all-ones key stays the largest
        HIP_TRY(rocprim::radix_sort_keys(sort_tmp, tmp_bytes, fwd, fout, (unsigned int)n, 0u, bits, st));
        HIP_TRY(rocprim::radix_sort_keys(sort_tmp, tmp_bytes, rev, rout, (unsigned int)n, 0u, bits, st));
    }
    mix_starts_kernel<<<gen_blocks((uint64_t)G + 1), GEN_BLOCK, 0, st>>>(fout, n, G, b, dstart);
    std::vector<uint64_t> start((size_t)G + 1);
    HIP_TRY(hipMemcpyAsync(start.data(), dstart, sizeof(uint64_t) * (G + 1), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const uint64_t A = start[G];
    std::vector<uint32_t> E(G);
    for (uint32_t i = 0; i < G; i++) E[i] = (uint32_t)(start[i + 1] - start[i]);
    if ((s = B.get(&Fc, A)) || (s = B.get(&Bc, A))) return s;
    mix_rows_kernel<<<gen_blocks(roff[G]), GEN_BLOCK, 0, st>>>(fout, dstart, droff, G, b, Fr);
    mix_cols_kernel<<<gen_blocks(A), GEN_BLOCK, 0, st>>>(fout, A, b, Fc);
    // the reverse keys' graph starts are the same (each graph has the same arcs)
    mix_rows_kernel<<<gen_blocks(roff[G]), GEN_BLOCK, 0, st>>>(rout, dstart, droff, G, b, Br);
    mix_cols_kernel<<<gen_blocks(A), GEN_BLOCK, 0, st>>>(rout, A, b, Bc);
    if (!ctx->batch) ctx->batch.reset(new (std::nothrow) BatchEngine());
    if (!ctx->batch) return fail(SCC_ERR_NOMEM, "host allocation failed");
    s = ctx->batch->upload_device(ctx->eng.device, st, G, V.data(), E.data(), roff.data(), start.data(), droff, Fr,
                                  Br, Fc, Bc);
    if (s != SCC_OK || !copy) return s;
    // a host copy in the generators' layout (for checking against them)
    std::memset(copy, 0, sizeof *copy);
    copy->num_graphs = G;
    copy->V = xalloc<uint32_t>(G);
    copy->E = xalloc<uint32_t>(G);
    copy->row_off = xalloc<uint64_t>((size_t)G + 1);
    copy->col_off = xalloc<uint64_t>((size_t)G + 1);
    copy->Fr = xalloc<uint32_t>(roff[G]);
    copy->Br = xalloc<uint32_t>(roff[G]);
    copy->Fc = xalloc<uint32_t>(A);
    copy->Bc = xalloc<uint32_t>(A);
    if (!copy->V || !copy->E || !copy->row_off || !copy->col_off || !copy->Fr || !copy->Br || !copy->Fc ||
        !copy->Bc) {
        scc_batch_host_free(copy);
        return fail(SCC_ERR_NOMEM, "host allocation of the batch copy failed");
    }
    std::memcpy(copy->V, V.data(), sizeof(uint32_t) * G);
    std::memcpy(copy->E, E.data(), sizeof(uint32_t) * G);
    std::memcpy(copy->row_off, roff.data(), sizeof(uint64_t) * (G + 1));
    std::memcpy(copy->col_off, start.data(), sizeof(uint64_t) * (G + 1));
    HIP_TRY(hipMemcpyAsync(copy->Fr, Fr, sizeof(uint32_t) * roff[G], hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(copy->Br, Br, sizeof(uint32_t) * roff[G], hipMemcpyDeviceToHost, st));
    if (A) {
        HIP_TRY(hipMemcpyAsync(copy->Fc, Fc, sizeof(uint32_t) * A, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(copy->Bc, Bc, sizeof(uint32_t) * A, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    return SCC_OK;
}

scc_status scc_graph_size(scc_ctx *ctx, uint32_t *V, uint32_t *E)
{
    CTX_DEV(ctx);
    if (!ctx->eng.has_graph) return fail(SCC_ERR_NOTREADY, "no graph uploaded");
    if (V) *V = ctx->eng.g.V;
    if (E) *E = ctx->eng.g.E;
    return SCC_OK;
}

scc_status scc_graph_download(scc_ctx *ctx, scc_host_graph *out)
{
    CTX_DEV(ctx);
    if (!out) return fail(SCC_ERR_INVALID, "null output");
    Engine &e = ctx->eng;
    if (!e.has_graph) return fail(SCC_ERR_NOTREADY, "no graph uploaded");
    std::memset(out, 0, sizeof *out);
    const size_t nr = (size_t)e.g.V + 2, nc = e.g.E;
    out->V = e.g.V;
    out->E = e.g.E;
    out->Fr = xalloc<uint32_t>(nr);
    out->Br = xalloc<uint32_t>(nr);
    out->Fc = xalloc<uint32_t>(nc);
    out->Bc = xalloc<uint32_t>(nc);
    if (!out->Fr || !out->Br || !out->Fc || !out->Bc) {
        scc_host_graph_free(out);
        return fail(SCC_ERR_NOMEM, "host allocation failed");
    }
    HIP_TRY(hipMemcpyAsync(out->Fr, e.g.Fr, sizeof(uint32_t) * nr, hipMemcpyDeviceToHost, e.stream));
    HIP_TRY(hipMemcpyAsync(out->Br, e.g.Br, sizeof(uint32_t) * nr, hipMemcpyDeviceToHost, e.stream));
    if (nc) {
        HIP_TRY(hipMemcpyAsync(out->Fc, e.g.Fc, sizeof(uint32_t) * nc, hipMemcpyDeviceToHost, e.stream));
        HIP_TRY(hipMemcpyAsync(out->Bc, e.g.Bc, sizeof(uint32_t) * nc, hipMemcpyDeviceToHost, e.stream));
    }
    HIP_TRY(hipStreamSynchronize(e.stream));
    return SCC_OK;
}

} // extern "C"
