// lds_queue.h -- workgroup-staged append to a global list.
//
// Same-address atomics serialise (about 10 ns each on MI355X), so a kernel
// that appends many items stages them in LDS and reserves global space
// with one atomic per flush (one per ~QBUF items) instead of one per wave.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sccamd {

constexpr uint32_t QBUF = 2048;

struct LdsQueue {
    uint32_t *buf;   // QBUF entries
    uint32_t *n;     // entries held
    uint32_t *base;  // broadcast of the reserved global offset
};

// Wave-aggregated append to the workgroup's LDS buffer.  The caller
// guarantees room for one push per lane since the last flush check.
__device__ __forceinline__ void lq_push(const LdsQueue &q, bool p, uint32_t v)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t act = __ballot(1);
    const uint64_t m = __ballot(p);
    if (!m) return;
    const int leader = __ffsll((unsigned long long)act) - 1;
    uint32_t b = 0;
    if ((int)lane == leader) b = atomicAdd(q.n, (uint32_t)__popcll(m));
    b = __shfl(b, leader);
    if (p) q.buf[b + (uint32_t)__popcll(m & ((1ull << lane) - 1))] = v;
}

// Flush when `force` or when another round of blockDim pushes might not
// fit.  Must be reached by every thread of the workgroup.
__device__ __forceinline__ void lq_flush(const LdsQueue &q, uint32_t *gcnt, uint32_t *gout, bool force)
{
    __syncthreads();
    const uint32_t n = *q.n;
    if (n == 0 || (!force && n + blockDim.x <= QBUF)) return;
    if (threadIdx.x == 0) *q.base = atomicAdd(gcnt, n);
    __syncthreads();
    const uint32_t b = *q.base;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) gout[b + i] = q.buf[i];
    __syncthreads();
    if (threadIdx.x == 0) *q.n = 0;
    __syncthreads();
}

// ---- wave-private queues ----
// A kernel whose waves should not wait on each other keeps one LDS queue
// per wave (nq is wave-uniform), reserves global space itself when its
// queue fills, and sends the remainders out with one reservation per
// 256-lane workgroup at the end.

// LDS written by some lanes of a wave and read by others of the same wave
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Append v for every lane with p (all 64 lanes call).
__device__ __forceinline__ void wave_push(uint32_t *wq, uint32_t &nq, bool p, uint32_t v)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t m = __ballot(p);
    if (p) wq[nq + (uint32_t)__popcll(m & ((1ull << lane) - 1))] = v;
    nq += (uint32_t)__popcll(m);
}

// Entries at or past `cap` are not written, and a reservation that passes
// it counts one in *anom (a caller whose bound says that cannot happen
// passes its region's size; the region's reader clamps its count to cap).
__device__ __forceinline__ void wave_flush(uint32_t *wq, uint32_t &nq, uint32_t *gcnt, uint32_t *gout,
                                           uint32_t cap = ~0u, uint32_t *anom = nullptr)
{
    if (!nq) return;
    const uint32_t lane = threadIdx.x & 63u;
    wave_sync();
    uint32_t b = 0;
    if (lane == 0) {
        b = atomicAdd(gcnt, nq);
        if (anom && b + nq > cap) atomicAdd(anom, 1u);
    }
    b = __shfl(b, 0);
    for (uint32_t i = lane; i < nq; i += 64)
        if (b + i < cap) gout[b + i] = wq[i];
    nq = 0;
    wave_sync();
}

// End of kernel: the four waves' remainders with one atomic (every thread
// of the 256-lane workgroup calls).  cap and anom as for wave_flush.
__device__ __forceinline__ void wave_flush_block(const uint32_t *wq, uint32_t nq, uint32_t *gcnt, uint32_t *gout,
                                                 uint32_t cap = ~0u, uint32_t *anom = nullptr)
{
    __shared__ uint32_t s_n[4], s_base;
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    if (lane == 0) s_n[wv] = nq;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t tot = s_n[0] + s_n[1] + s_n[2] + s_n[3];
        s_base = tot ? atomicAdd(gcnt, tot) : 0u;
        if (anom && tot && s_base + tot > cap) atomicAdd(anom, 1u);
    }
    __syncthreads();
    uint32_t off = s_base;
    for (uint32_t w = 0; w < wv; w++) off += s_n[w];
    for (uint32_t i = lane; i < nq; i += 64)
        if (off + i < cap) gout[off + i] = wq[i];
}

} // namespace sccamd
