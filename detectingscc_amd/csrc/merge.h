// merge.h -- launchers of the partitioned local SCC and quotient-graph merge
// kernels (kernels_merge.hip) and the device exclusive scan they use.
#pragma once
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
void k_q_map(hipStream_t s, const DevGraph &g, DevState &st, const uint32_t *tbits, const uint32_t *qid,
             const uint32_t *qlab, const uint32_t *rep);

} // namespace sccamd
