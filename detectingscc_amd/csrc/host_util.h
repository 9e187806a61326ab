// host_util.h -- host-side helpers shared by the library's C++ sources:
// error reporting for the C ABI, a small thread pool parallel_for, and the
// counter-based random number generator used by the graph generators.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "scc_amd.h"

namespace sccamd {

// Records a message for scc_last_error() on this thread and returns `st`.
scc_status fail(scc_status st, const char *fmt, ...);
void clear_error();

// Number of host threads used by parallel loops: SCC_HOST_THREADS if set,
// otherwise the hardware concurrency capped at 16.
int host_threads();

// Runs f(begin, end, thread_index) over [0, n) split into contiguous
// chunks, one per thread.
void parallel_chunks(uint64_t n, const std::function<void(uint64_t, uint64_t, int)> &f,
                     int nthreads = 0);

// splitmix64 finaliser; rnd(seed, stream, i) is a pure function of its
// arguments, so generated graphs do not depend on the thread count.
inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

inline uint64_t rnd(uint64_t seed, uint64_t stream, uint64_t i)
{
    return mix64(mix64(seed + stream * 0x9E3779B97F4A7C15ull) + i * 0xD1B54A32D192ED03ull);
}

// Uniform integer in [0, n) from a 64-bit random value (multiply-shift).
inline uint64_t below(uint64_t r, uint64_t n)
{
    return (uint64_t)(((unsigned __int128)r * n) >> 64);
}

// Uniform double in [0, 1) from a 64-bit random value.
inline double unit(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

// True with probability p for a uniform 64-bit r.
inline bool bernoulli(uint64_t r, double p) { return p >= 1.0 || unit(r) < p; }

template <class T> T *xalloc(size_t n)
{
    return static_cast<T *>(std::malloc(sizeof(T) * (n ? n : 1)));
}

} // namespace sccamd
