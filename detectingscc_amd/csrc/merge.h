// merge.h -- launchers of the partitioned local SCC and quotient-graph merge
// kernels (kernels_merge.hip) and the device exclusive scan they use.
#pragma once
#include "scc_amd.h"
#include "scc_dev.h"

namespace sccamd {

// Exclusive prefix sum of n u32 (out may alias in); tile_sums needs
// scan_tile_count(n) words.
void k_exscan(hipStream_t s, const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *tile_sums);
uint64_t scan_tile_count(uint64_t n);

// live vertices -> class 3*part[v]; tbits = finished set before the local phase
void k_local_init(hipStream_t s, const DevGraph &g, DevState &st, uint32_t *tbits);
void k_q_mark(hipStream_t s, const DevGraph &g, const DevState &st, const uint32_t *tbits, uint32_t *isrep);
void k_q_arcs(hipStream_t s, int pass, const DevGraph &g, const DevState &st, const uint32_t *tbits,
              const uint32_t *qid, uint32_t *cnt_out, uint32_t *cnt_in, uint32_t *Qc, uint32_t *QBc);
void k_q_rep(hipStream_t s, const DevGraph &g, const uint32_t *isrep, const uint32_t *qid, uint32_t *rep);
// qv[v] = 1 + quotient id of v's local SCC, 0 for a vertex finished before
// the local phase (V+1 words)
void k_q_vid(hipStream_t s, const DevGraph &g, const DevState &st, const uint32_t *tbits, const uint32_t *qid,
             uint32_t *qv);
// Arcs between different local SCCs: counted into *n (emit false; *n must
// start at 0), or appended to keys as (source << 32 | target) quotient ids
// 1..QV (emit true; *n must start at 0 and ends at the count; entries at or
// past cap are not written).
void k_q_keys(hipStream_t s, bool emit, const DevGraph &g, const uint32_t *qv, unsigned long long *n,
              uint64_t *keys, uint64_t cap);
// The CSR (R, C) and reverse CSR (RB, CB) of V vertices from n arc keys
// (source << 32 | target, ids 1..V) in `keys`, by two radix sorts with
// `out` (n words) as the other buffer; `keys` is overwritten.  `tmp` holds
// csr_sort_bytes(V, n) bytes.
size_t csr_sort_bytes(uint32_t V, uint64_t n);
scc_status csr_from_keys(hipStream_t st, uint32_t V, uint64_t n, uint64_t *keys, uint64_t *out, void *tmp,
                         size_t tmp_bytes, uint32_t *R, uint32_t *C, uint32_t *RB, uint32_t *CB);
void k_q_map(hipStream_t s, const DevGraph &g, DevState &st, const uint32_t *tbits, const uint32_t *qid,
             const uint32_t *qlab, const uint32_t *rep);

} // namespace sccamd
