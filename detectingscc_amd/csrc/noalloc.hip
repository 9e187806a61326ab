// noalloc.hip -- the allocation-free variant of scc_histogram.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) get the size
// histogram of the last run through scc_histogram_noalloc, whose staging
// (the compacted (size, count) pairs and their counter) lives in the
// file-scope device array `scratch` below instead of a buffer the library
// allocates on first use.  The per-size counts use the context's own label
// and state arrays, which scc_graph_upload allocated.  Each GPU holds one
// copy of `scratch` (262,400 bytes) until the process ends; every call on a
// GPU shares it, so calls that overlap on one GPU race.
#include <algorithm>
#include <mutex>

#include "engine.h"
#include "host_util.h"

namespace sccamd {

// Largest number of distinct SCC sizes the array holds.  Distinct sizes of
// disjoint SCCs sum to at most V, so a graph of V vertices has at most k
// distinct sizes with k (k + 1) / 2 <= V: 16,384 covers every graph of up
// to 134,225,920 vertices.
constexpr uint32_t kNoAllocPairs = 16384;
// the pairs (two 64-bit words each), then one 256-byte line for the counter
constexpr size_t kNoAllocWords = 2 * (size_t)kNoAllocPairs + 32;

static __device__ __align__(256) uint64_t scratch[kNoAllocWords];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint64_t *g_scratch[kMaxDevices];   // device address of `scratch` on each GPU

// The device address of this GPU's copy of `scratch`, looked up once per
// device at context creation with that device current.  In host code the
// array's name is only a stand-in; kernels and copies use this address.
scc_status noalloc_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(SCC_ERR_INVALID, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return SCC_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return fail(SCC_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                    hipGetErrorString(e));
    }
    g_scratch[device] = static_cast<uint64_t *>(p);
    return SCC_OK;
}

// Sorts the first min(*npairs, cap) (size, count) pairs of `pairs` by size
// in one workgroup: packed as size << 32 | count (both below 2^32, sizes
// distinct) into LDS, bitonic sort over the next power of two, unpacked.
constexpr uint32_t kSortThreads = 1024;

__global__ __launch_bounds__(kSortThreads) void pairs_sort_kernel(uint64_t *pairs, const uint32_t *npairs,
                                                                  uint32_t cap)
{
    __shared__ uint64_t s[kNoAllocPairs];
    const uint32_t n = min(*npairs, cap);
    uint32_t N = 1;
    while (N < n) N <<= 1;
    for (uint32_t i = threadIdx.x; i < N; i += kSortThreads)
        s[i] = i < n ? (pairs[2 * (size_t)i] << 32) | (pairs[2 * (size_t)i + 1] & 0xFFFFFFFFull) : ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= N; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < N; i += kSortThreads) {
                const uint32_t x = i ^ j;
                if (x > i) {
                    const uint64_t a = s[i], b = s[x];
                    if ((a > b) == ((i & k) == 0)) {
                        s[i] = b;
                        s[x] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = threadIdx.x; i < n; i += kSortThreads) {
        pairs[2 * (size_t)i] = s[i] >> 32;
        pairs[2 * (size_t)i + 1] = s[i] & 0xFFFFFFFFull;
    }
}

} // namespace sccamd

using namespace sccamd;

#define HIP_TRY(call)                                                                       \
    do {                                                                                    \
        hipError_t _e = (call);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail(SCC_ERR_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), \
                        __FILE__, __LINE__);                                                \
    } while (0)

extern "C" scc_status scc_histogram_noalloc(scc_ctx *ctx, uint64_t *pairs, uint32_t cap, uint32_t *npairs)
{
    clear_error();
    if (!ctx) return fail(SCC_ERR_INVALID, "null context");
    if (!npairs || (cap && !pairs)) return fail(SCC_ERR_INVALID, "null output");
    Engine &E = ctx->eng;
    if (!E.ran) return fail(SCC_ERR_NOTREADY, "no completed run");
    uint64_t kmax = 1;
    while (kmax * (kmax + 1) / 2 <= E.g.V) kmax++;
    if (kmax > kNoAllocPairs)
        return fail(SCC_ERR_INVALID,
                    "V=%u may have up to %llu distinct SCC sizes; the allocation-free histogram holds %u "
                    "(graphs of up to 134,225,920 vertices)",
                    E.g.V, (unsigned long long)kmax, kNoAllocPairs);
    if (E.device < 0 || E.device >= kMaxDevices) return fail(SCC_ERR_INVALID, "bad device");
    uint64_t *d;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        d = g_scratch[E.device];
    }
    if (!d) return fail(SCC_ERR_NOTREADY, "scratch address of device %d not set up", E.device);
    HIP_TRY(hipSetDevice(E.device));
    uint32_t *dn = reinterpret_cast<uint32_t *>(d + 2 * (size_t)kNoAllocPairs);
    k_fill_u32(E.stream, dn, 0u, 1);   // a fill kernel, not a memset node
    k_sizes(E.stream, E.g, E.st);
    k_size_histogram(E.stream, E.g, E.st, E.st.aux, d, kNoAllocPairs, dn);
    pairs_sort_kernel<<<1, kSortThreads, 0, E.stream>>>(d, dn, kNoAllocPairs);
    uint32_t n = 0;
    HIP_TRY(hipMemcpyAsync(&n, dn, sizeof n, hipMemcpyDeviceToHost, E.stream));
    HIP_TRY(hipStreamSynchronize(E.stream));
    if (n > kNoAllocPairs) return fail(SCC_ERR_DEVICE, "%u distinct SCC sizes exceed %u", n, kNoAllocPairs);
    const uint32_t out = std::min(n, cap);
    if (out) {
        HIP_TRY(hipMemcpyAsync(pairs, d, 2 * sizeof(uint64_t) * out, hipMemcpyDeviceToHost, E.stream));
        HIP_TRY(hipStreamSynchronize(E.stream));
    }
    *npairs = n;
    return SCC_OK;
}
