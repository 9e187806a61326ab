// batch.h -- device layout and launcher of the batched persistent SCC kernel.
#pragma once
#include "scc_dev.h"

namespace sccamd {

// Everything the persistent kernel reads, per batch (see kernels_batch.hip
// for the slot layout).
struct BatchDev {
    uint32_t G = 0;
    const uint64_t *voff = nullptr, *eoff = nullptr;   // G+1 each
    const uint32_t *Vn = nullptr;                      // G
    const uint32_t *Fr = nullptr, *Fc = nullptr, *Br = nullptr, *Bc = nullptr;
    uint32_t *cls = nullptr, *lab = nullptr, *aux = nullptr;
    uint32_t *fvis = nullptr, *bvis = nullptr;        // bit voff[i]+v
    uint64_t *piv = nullptr;                           // 3*voff[i] + class
    uint32_t *q = nullptr;                             // 4*voff[i]: FW in/out, BW in/out
    uint32_t *pivlist = nullptr;                       // voff[i]
    const uint32_t *cfg = nullptr;                     // per graph
    const uint32_t *order = nullptr;                   // processing order (indices)
    uint32_t *nscc = nullptr;                          // per graph
    uint64_t *tstart = nullptr, *tend = nullptr;       // per graph wall-clock ticks
    uint32_t *next = nullptr;                          // work-queue head
};

// nscc value of a graph whose run hit an internal inconsistency
constexpr uint32_t BATCH_ANOMALY = 0xFFFFFFFFu;

// cfg word of one graph
uint32_t batch_cfg(int lanes, bool trim1, bool trim2, bool slota);

// Persistent launch over order[first .. first+count) with `block` lanes per
// workgroup (256, 512 or 1024) and `grid` workgroups; *B.next must be 0.
void k_batch(hipStream_t s, const BatchDev &B, uint32_t first, uint32_t count, int block, uint32_t grid);
// workgroups per CU that fit at the given block size
int batch_occupancy(int block);

} // namespace sccamd
