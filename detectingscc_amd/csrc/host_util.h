// host_util.h -- host-side helpers shared by the library's C++ sources:
// error reporting for the C ABI and a small thread pool parallel_for (the
// generators' random numbers are in rng.h).
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

#include "rng.h"
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

template <class T> T *xalloc(size_t n)
{
    return static_cast<T *>(std::malloc(sizeof(T) * (n ? n : 1)));
}

} // namespace sccamd
