// kernels_reach.hip -- one BFS level of forward / backward reachability
// (fwd/bwd scc_kernels.cu:360-445, fwd_warp/bwd_warp scc_kernels.h:46-105)
// rebuilt as data-driven frontier expansion for gfx950.
//
// A level reads the input queue of each direction and writes the next one.
// Claims are atomicOr on a visited bitmap (finished vertices pre-set), so a
// vertex enters a queue at most once per search; pushes are aggregated per
// wave with a ballot, a popcount prefix and one atomicAdd.  Queue counts
// rotate through three device words so that no host step sits between
// levels: level L reads cnt[L%3], appends to cnt[(L+1)%3] and zeroes
// cnt[(L+2)%3] for level L+1.  With `single_class` every live vertex is in
// one class and the class compare is skipped (the probe is one bitmap bit).
#include "scc_dev.h"
#include "scc_amd.h"

namespace sccamd {

__device__ __forceinline__ uint32_t wlane() { return threadIdx.x & 63u; }

template <bool SINGLE>
__device__ __forceinline__ bool try_claim(uint32_t *vis, const uint32_t *cls, uint32_t w, uint32_t c)
{
    const uint32_t bit = 1u << (w & 31);
    if (vis[w >> 5] & bit) return false;
    if (!SINGLE && cls[w] != c) return false;
    return !(atomicOr(&vis[w >> 5], bit) & bit);
}

// Append w for every active lane with `claimed`; one atomic per wave.
__device__ __forceinline__ void push_wave(bool claimed, uint32_t w, uint32_t *qout, uint32_t *cnt)
{
    const uint64_t act = __ballot(1);
    const uint64_t m = __ballot(claimed);
    if (!m) return;
    const int leader = __ffsll((unsigned long long)act) - 1;
    uint32_t base = 0;
    if ((int)wlane() == leader) base = atomicAdd(cnt, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (claimed) qout[base + (uint32_t)__popcll(m & ((1ull << wlane()) - 1))] = w;
}

// W lanes per frontier vertex; W = 1 is the vertex-per-thread mapping of
// fwd/bwd, W > 1 the virtual warp of fwd_warp<W> (lanes stride the list).
template <int W, bool SINGLE>
__global__ __launch_bounds__(256) void expand_vw_kernel(LevelArgs a, uint32_t bpd)
{
    const int dir = (a.ndirs == 2 && blockIdx.x >= bpd) ? 1 : 0;
    const uint32_t bid = blockIdx.x - (uint32_t)dir * bpd;
    const DirArgs &D = a.d[dir];
    const int L = a.level;
    const uint32_t n = D.cnt[L % 3];
    uint32_t *cnt_out = &D.cnt[(L + 1) % 3];
    if (bid == 0 && threadIdx.x == 0) D.cnt[(L + 2) % 3] = 0;
    if (n == 0) return;
    const uint32_t g = (bid * blockDim.x + threadIdx.x) / W;
    const uint32_t ng = bpd * blockDim.x / W;
    const uint32_t sub = threadIdx.x % W;
    const uint32_t *__restrict__ R = D.R;
    const uint32_t *__restrict__ C = D.C;
    for (uint32_t i = g; i < n; i += ng) {
        const uint32_t v = D.qin[i];
        const uint32_t c = SINGLE ? 0u : a.cls[v];
        const uint32_t e = R[v + 1];
        for (uint32_t j = R[v] + sub; j < e; j += W) {
            const uint32_t w = C[j];
            const bool claimed = try_claim<SINGLE>(D.vis, a.cls, w, c);
            push_wave(claimed, w, D.qout, cnt_out);
        }
    }
}

template <int W>
static void launch_vw(hipStream_t s, const LevelArgs &a, bool single, uint32_t bpd)
{
    const uint32_t grid = bpd * (uint32_t)a.ndirs;
    if (single)
        expand_vw_kernel<W, true><<<grid, 256, 0, s>>>(a, bpd);
    else
        expand_vw_kernel<W, false><<<grid, 256, 0, s>>>(a, bpd);
}

void k_expand_level(hipStream_t s, const LevelArgs &a, int expand, int warp, bool single_class,
                    uint32_t blocks_per_dir)
{
    int W = expand == SCC_EXPAND_VWARP ? warp : 1;
    switch (W) {
    case 1: launch_vw<1>(s, a, single_class, blocks_per_dir); break;
    case 2: launch_vw<2>(s, a, single_class, blocks_per_dir); break;
    case 4: launch_vw<4>(s, a, single_class, blocks_per_dir); break;
    case 8: launch_vw<8>(s, a, single_class, blocks_per_dir); break;
    case 16: launch_vw<16>(s, a, single_class, blocks_per_dir); break;
    case 32: launch_vw<32>(s, a, single_class, blocks_per_dir); break;
    default: launch_vw<64>(s, a, single_class, blocks_per_dir); break;
    }
}

} // namespace sccamd
