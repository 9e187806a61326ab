// engine.hip -- the SCC phase engine: device-resident graph and state, the
// vHong / wHong / vSlota / wSlota phase orders, phase timing, and the C ABI
// for the device engine.
//
// The reference drives every phase from the host with a 1-byte D2H after
// each kernel (scc_WCC.cu:683-686).  Here queue counts and fixpoint flags
// live on the device: a direction-optimised search decides each level on
// the device and the host queues windows of levels per look at the
// counters; the reference mappings queue `check_every` levels; trim-1
// queues several passes per look; other fixpoints look once per sweep.
// A look is a one-workgroup kernel that publishes the counters into
// mapped host memory, on which the host spins.
#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include <algorithm>

#include "engine.h"
#include "host_util.h"
#include "merge.h"

namespace sccamd {

#define HIP_TRY(call)                                                                       \
    do {                                                                                    \
        hipError_t _e = (call);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail(SCC_ERR_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), \
                        __FILE__, __LINE__);                                                \
    } while (0)

static const char *const kPhaseNames[SCC_PHASE_COUNT] = {
    "trim1", "trim2", "pivot", "reach", "update", "wcc", "color", "local", "merge", "label"};

const char *phase_name(int p) { return (p >= 0 && p < SCC_PHASE_COUNT) ? kPhaseNames[p] : "?"; }

Engine::~Engine() { release(); }

void Engine::release()
{
    if (device >= 0) (void)hipSetDevice(device);
    auto f = [](void *p) {
        if (p) (void)hipFree(p);
    };
    f(g.Fr); f(g.Fc); f(g.Br); f(g.Bc); f(g.F1); f(g.B1); f(g.DP);
    f(st.cls); f(st.lab); f(st.aux); f(st.fvis); f(st.bvis); f(st.piv);
    for (int i = 0; i < 2; i++) {
        f(st.qf[i]); f(st.qb[i]);
    }
    for (int i = 0; i < 3; i++) {
        f(st.bigf[i]); f(st.bigb[i]); f(st.bigoff_f[i]); f(st.bigoff_b[i]);
    }
    f(st.pivlist); f(st.stamp); f(st.part); f(st.ctr); f(st.bred); f(d_tmp8);
    for (auto &b : buf) {
        f(b.p);
        b = DevBuf();
    }
    sub.reset();
    if (h_ctr) (void)hipHostFree(h_ctr);
    if (h_pub) (void)hipHostFree(h_pub);
    h_pub = nullptr;
    d_pub = nullptr;
    for (auto e : events) (void)hipEventDestroy(e);
    events.clear();
    if (stream) (void)hipStreamDestroy(stream);
    g = DevGraph();
    st = DevState();
    d_tmp8 = nullptr;
    h_ctr = nullptr;
    stream = nullptr;
    capV = capE = 0;
}

scc_status Engine::init(int dev)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return fail(SCC_ERR_DEVICE, "no HIP device available");
    if (dev < 0 || dev >= n) return fail(SCC_ERR_INVALID, "device %d out of range (0..%d)", dev, n - 1);
    device = dev;
    HIP_TRY(hipSetDevice(dev));
    scc_status ns = noalloc_init(dev);
    if (ns != SCC_OK) return ns;
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIP_TRY(hipHostMalloc((void **)&h_ctr, CT(kCtrLines) * sizeof(uint32_t), hipHostMallocDefault));
    HIP_TRY(hipMalloc((void **)&st.ctr, CT(kCtrLines) * sizeof(uint32_t)));
    HIP_TRY(hipMemset(st.ctr, 0, CT(kCtrLines) * sizeof(uint32_t)));
    // counters published into mapped, coherent host memory: C_COUNT lines
    // plus one line for the sequence word (without it, read-backs copy and
    // synchronise the stream)
    if (hipHostMalloc((void **)&h_pub, 2 * CT(C_COUNT + 1) * sizeof(uint32_t),
                      hipHostMallocMapped | hipHostMallocCoherent) == hipSuccess) {
        std::memset(h_pub, 0, 2 * CT(C_COUNT + 1) * sizeof(uint32_t));
        if (hipHostGetDevicePointer((void **)&d_pub, h_pub, 0) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipHostFree(h_pub);
            h_pub = nullptr;
            d_pub = nullptr;
        }
    } else {
        (void)hipGetLastError();
        h_pub = nullptr;
    }
    // diagnostic: SCC_TRACE_LEVELS=1 prints every host decision of a search
    const char *tr = std::getenv("SCC_TRACE_LEVELS");
    trace_levels = tr && tr[0] == '1';
    return SCC_OK;
}

template <class T> static hipError_t dalloc(T **p, size_t n)
{
    return hipMalloc((void **)p, sizeof(T) * (n ? n : 1));
}

// (Re)allocate device arrays for V vertices and E arcs.
scc_status Engine::reserve(uint32_t V, uint32_t E)
{
    HIP_TRY(hipSetDevice(device));
    // the queues' slot regions depend on V through the grid sizes: a smaller
    // graph in a larger context must fit the queues as allocated (they grow
    // with V, and this keeps it so)
    const bool queues_fit = std::max(queue_words(V), trim_list_words(V)) <= qf_words && queue_words(V) <= qb_words;
    if (V <= capV && E <= capE && queues_fit) return SCC_OK;
    uint32_t nV = V > capV ? V : capV, nE = E > capE ? E : capE;
    // keep the context, drop the arrays
    auto f = [](void *p) {
        if (p) (void)hipFree(p);
    };
    f(g.Fr); f(g.Fc); f(g.Br); f(g.Bc); f(g.F1); f(g.B1); f(g.DP);
    f(st.cls); f(st.lab); f(st.aux); f(st.fvis); f(st.bvis); f(st.piv);
    for (int i = 0; i < 2; i++) {
        f(st.qf[i]); f(st.qb[i]);
    }
    for (int i = 0; i < 3; i++) {
        f(st.bigf[i]); f(st.bigb[i]); f(st.bigoff_f[i]); f(st.bigoff_b[i]);
    }
    f(st.pivlist); f(st.stamp); f(st.part); f(st.bred); f(d_tmp8);
    uint32_t *ctr = st.ctr;
    g = DevGraph();
    st = DevState();
    st.ctr = ctr;
    capV = capE = 0;
    const size_t nv = (size_t)nV + 1;
    const size_t nwords = 2 * ((nv + 63) / 64);
    hipError_t e = hipSuccess;
    auto A = [&](hipError_t r) {
        if (r != hipSuccess && e == hipSuccess) e = r;
    };
    A(dalloc(&g.Fr, nv + 1)); A(dalloc(&g.Br, nv + 1));
    A(dalloc(&g.Fc, nE)); A(dalloc(&g.Bc, nE));
    A(dalloc(&g.F1, nv)); A(dalloc(&g.B1, nv)); A(dalloc(&g.DP, nv));
    A(dalloc(&st.cls, nv)); A(dalloc(&st.lab, nv)); A(dalloc(&st.aux, nv));
    A(dalloc(&st.fvis, 2 * nwords)); A(dalloc(&st.bvis, 2 * nwords));
    A(dalloc(&st.piv, 3 * nv + 3));
    // the FW queues also hold the trim-1 lists; sized for the largest of
    // nV and V (slot regions follow the grid sizes of the graph uploaded)
    qf_words = std::max({queue_words(nV), trim_list_words(nV), queue_words(V), trim_list_words(V)});
    qb_words = std::max(queue_words(nV), queue_words(V));
    for (int i = 0; i < 2; i++) {
        A(dalloc(&st.qf[i], qf_words)); A(dalloc(&st.qb[i], qb_words));
    }
    // a hub list holds vertices of degree >= LB_BIG only
    const size_t nhub = std::min<size_t>(nv, (size_t)nE / LB_BIG + 1);
    for (int i = 0; i < 3; i++) {
        A(dalloc(&st.bigf[i], nhub)); A(dalloc(&st.bigb[i], nhub));
        A(dalloc(&st.bigoff_f[i], nhub)); A(dalloc(&st.bigoff_b[i], nhub));
    }
    // d_tmp8 also holds the single-class trim's live bitmap (live_bits())
    A(dalloc(&st.pivlist, nv)); A(dalloc(&st.stamp, nv)); A(dalloc(&st.part, nv + 1));
    A(dalloc(&d_tmp8, std::max(nv, 4 * (nwords + 2))));
    A(dalloc(&st.bred, 2 * (size_t)VIS_BLOCKS));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(e == hipErrorOutOfMemory ? SCC_ERR_NOMEM : SCC_ERR_DEVICE,
                    "Error: Not enough memory on GPU (V=%u E=%u): %s", nV, nE, hipGetErrorString(e));
    }
    st.nwords = (uint32_t)nwords;
    capV = nV;
    capE = nE;
    k_fill_u32(stream, st.stamp, 0u, nv);   // no epoch is 0
    trim_epoch = 0;
    return SCC_OK;
}

// Host-side validation of the reference layout: a malformed CSR must be
// rejected before any kernel indexes with it.
static scc_status validate_csr(uint32_t V, uint32_t E, const uint32_t *R, const uint32_t *C,
                               const char *name)
{
    if (R[0] != 0 || R[1] != 0 || R[V + 1] != E)
        return fail(SCC_ERR_INVALID, "%s row array must start 0,0 and end at E=%u", name, E);
    std::atomic<int> bad{0};
    parallel_chunks((uint64_t)V + 1, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t v = b; v < e; v++)
            if (R[v] > R[v + 1]) bad = 1;
    });
    if (bad) return fail(SCC_ERR_INVALID, "%s row array is not monotone", name);
    parallel_chunks(E, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t i = b; i < e; i++)
            if (C[i] < 1 || C[i] > V) bad = 1;
    });
    if (bad) return fail(SCC_ERR_INVALID, "%s column entry outside [1, %u]", name, V);
    return SCC_OK;
}

scc_status Engine::upload(uint32_t V, uint32_t E, const uint32_t *Fr, const uint32_t *Fc,
                          const uint32_t *Br, const uint32_t *Bc, bool checked)
{
    if (!Fr || !Br || (E && (!Fc || !Bc))) return fail(SCC_ERR_INVALID, "null graph array");
    if (V >= (1u << 30)) return fail(SCC_ERR_INVALID, "V=%u exceeds 2^30-1", V);
    scc_status s;
    if (!checked && (s = validate_csr(V, E, Fr, Fc, "Fr/Fc")) != SCC_OK) return s;
    if (!checked && (s = validate_csr(V, E, Br, Bc, "Br/Bc")) != SCC_OK) return s;
    if ((s = reserve(V, E)) != SCC_OK) return s;
    g.V = V;
    g.E = E;
    HIP_TRY(hipMemcpyAsync(g.Fr, Fr, sizeof(uint32_t) * ((size_t)V + 2), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(g.Br, Br, sizeof(uint32_t) * ((size_t)V + 2), hipMemcpyHostToDevice, stream));
    if (E) {
        HIP_TRY(hipMemcpyAsync(g.Fc, Fc, sizeof(uint32_t) * E, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(g.Bc, Bc, sizeof(uint32_t) * E, hipMemcpyHostToDevice, stream));
    }
    HIP_TRY(hipStreamSynchronize(stream));
    has_graph = true;
    has_part = false;
    ran = false;
    return SCC_OK;
}

// Partition ids are remapped to dense 0..K-1 (order kept) so that the local
// classes 3*id stay inside the pivot table.
scc_status Engine::upload_partition(const uint32_t *part)
{
    if (!has_graph) return fail(SCC_ERR_NOTREADY, "no graph uploaded");
    std::vector<uint32_t> ids(part + 1, part + 1 + g.V);
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    std::vector<uint32_t> dense((size_t)g.V + 2, 0);
    parallel_chunks(g.V, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t v = b + 1; v <= e; v++)
            dense[v] = (uint32_t)(std::lower_bound(ids.begin(), ids.end(), part[v]) - ids.begin());
    });
    // on the engine's stream (non-blocking, so not ordered with the null
    // stream): after any queued work that reads the partition, before the
    // next run's; the synchronisation keeps `dense` alive until copied
    HIP_TRY(hipMemcpyAsync(st.part, dense.data(), sizeof(uint32_t) * ((size_t)g.V + 2), hipMemcpyHostToDevice,
                           stream));
    HIP_TRY(hipStreamSynchronize(stream));
    num_parts = (uint32_t)ids.size();
    has_part = true;
    return SCC_OK;
}

scc_status Engine::adopt_begin(uint32_t V, uint32_t E)
{
    if (V >= (1u << 30)) return fail(SCC_ERR_INVALID, "V=%u exceeds 2^30-1", V);
    has_graph = has_part = ran = false;
    scc_status s = reserve(V, E);
    if (s != SCC_OK) return s;
    g.V = V;
    g.E = E;
    return SCC_OK;
}

void Engine::adopt_end()
{
    has_graph = true;
    has_part = false;
    ran = false;
}

// Adopt a graph that is already in device memory (copied device to device).
// Used for the quotient graph, whose CSR is built by kernels and valid by
// construction.
scc_status Engine::upload_device(uint32_t V, uint32_t E, const uint32_t *dFr, const uint32_t *dFc,
                                 const uint32_t *dBr, const uint32_t *dBc, hipStream_t from)
{
    if (V >= (1u << 30)) return fail(SCC_ERR_INVALID, "V=%u exceeds 2^30-1", V);
    scc_status s;
    if ((s = reserve(V, E)) != SCC_OK) return s;
    g.V = V;
    g.E = E;
    HIP_TRY(hipStreamSynchronize(from));
    HIP_TRY(hipMemcpyAsync(g.Fr, dFr, sizeof(uint32_t) * ((size_t)V + 2), hipMemcpyDeviceToDevice, stream));
    HIP_TRY(hipMemcpyAsync(g.Br, dBr, sizeof(uint32_t) * ((size_t)V + 2), hipMemcpyDeviceToDevice, stream));
    if (E) {
        HIP_TRY(hipMemcpyAsync(g.Fc, dFc, sizeof(uint32_t) * E, hipMemcpyDeviceToDevice, stream));
        HIP_TRY(hipMemcpyAsync(g.Bc, dBc, sizeof(uint32_t) * E, hipMemcpyDeviceToDevice, stream));
    }
    has_graph = true;
    has_part = false;
    ran = false;
    return SCC_OK;
}

template <class T> scc_status Engine::scratch(DevBuf &b, size_t n, T **out)
{
    const size_t bytes = sizeof(T) * (n ? n : 1);
    if (b.cap < bytes) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        hipError_t e = hipMalloc(&b.p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(SCC_ERR_NOMEM, "Error: Not enough memory on GPU (%zu scratch bytes)", bytes);
        }
        b.cap = bytes;
    }
    *out = static_cast<T *>(b.p);
    return SCC_OK;
}

// Both words of every counter line into host memory the GPU writes
// directly, then the sequence number (system-scope release).
// Words of each counter line the host reads: 32- and 64-bit counters use
// words 0..1, the search state (DirState) words 0..kDirStateWords-1.
constexpr int kPubWords = kDirStateWords;

// A spread counter is published as its own word plus its slots.
__global__ void publish_ctr_kernel(const uint32_t *ctr, uint32_t *pub, uint32_t seq)
{
    const uint32_t i = threadIdx.x;
    if (i < C_COUNT) {
        for (int w = 0; w < kPubWords; w++) pub[CT(i) + w] = ctr[CT(i) + w];
        if (spread_index((int)i) >= 0) {
            uint32_t t = ctr[CT(i)];
            for (int j = 0; j < kSpread; j++) t += ctr[CSLOT((int)i, j)];
            pub[CT(i)] = t;
        }
    }
    __threadfence_system();
    __syncthreads();
    if (i == 0) __hip_atomic_store(pub + CT(C_COUNT), seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The host reads the counters of everything queued before a publication.
// With the published mirror it spins on the sequence word (a few
// microseconds after the last kernel) instead of a copy plus a stream
// synchronisation (about 25 us); after kPublishWaitMs it synchronises the
// stream anyway, which also reports any device error.  The mirror has two
// buffers, publication k going to buffer k & 1, so one publication may be
// queued behind the one the host waits for (pipelined search windows): a
// caller waits for publication k before it queues k + 2.
static constexpr double kPublishWaitMs = 2000.0;
static constexpr size_t kPubBuffer = CT(C_COUNT + 1);   // words of one buffer, sequence word last

uint32_t Engine::publish_ctr()
{
    const uint32_t seq = ++pub_seq;
    if (h_pub) publish_ctr_kernel<<<1, 64, 0, stream>>>(st.ctr, d_pub + (seq & 1) * kPubBuffer, seq);
    return seq;
}

scc_status Engine::wait_ctr(uint32_t seq)
{
    if (!h_pub) {
        HIP_TRY(hipMemcpyAsync(h_ctr, st.ctr, CT(kCtrLines) * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        for (int i = 0; i < C_COUNT; i++)
            if (spread_index(i) >= 0)
                for (int j = 0; j < kSpread; j++) h_ctr[CT(i)] += h_ctr[CSLOT(i, j)];
        return SCC_OK;
    }
    HIP_TRY(hipGetLastError());
    const uint32_t *buf = h_pub + (seq & 1) * kPubBuffer;
    const uint32_t *flag = buf + CT(C_COUNT);
    const auto t0 = std::chrono::steady_clock::now();
    for (uint64_t spin = 0; __atomic_load_n(flag, __ATOMIC_ACQUIRE) != seq; spin++) {
        if ((spin & 1023) == 1023 &&
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() > kPublishWaitMs) {
            HIP_TRY(hipStreamSynchronize(stream));
            if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) != seq)
                return fail(SCC_ERR_DEVICE, "counter publication %u did not arrive", seq);
            break;
        }
        __builtin_ia32_pause();
    }
    for (int i = 0; i < C_COUNT; i++)
        for (int w = 0; w < kPubWords; w++) h_ctr[CT(i) + w] = buf[CT(i) + w];
    return SCC_OK;
}

scc_status Engine::read_ctr() { return wait_ctr(publish_ctr()); }

// Zero counters [first, first+n), both words of each (64-bit counters too)
// and the slots of spread counters.
__global__ void zero_ctr_kernel(uint32_t *ctr, uint32_t first, uint32_t n)
{
    if (threadIdx.x < n) {
        const int c = (int)(first + threadIdx.x);
        ctr[CT(c)] = 0;
        ctr[CT(c) + 1] = 0;
        if (spread_index(c) >= 0)
            for (int j = 0; j < kSpread; j++) ctr[CSLOT(c, j)] = 0;
    }
}

void Engine::zero_ctr(int first, int n)
{
    zero_ctr_kernel<<<1, 64, 0, stream>>>(st.ctr, (uint32_t)first, (uint32_t)n);
}

void Engine::mark(int phase, bool force)
{
    if (phase >= 0 && !phase_marks && !force) return;
    if (nmarks == events.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        events.push_back(e);
    }
    (void)hipEventRecord(events[nmarks], stream);
    mark_phase.resize(nmarks + 1);
    mark_phase[nmarks] = phase;
    nmarks++;
}

// Phase times: the interval between consecutive marks belongs to the
// phase of the earlier mark (-1 = untimed).
void Engine::collect_times(scc_stats *s)
{
    for (int p = 0; p < SCC_PHASE_COUNT; p++) s->ms_phase[p] = 0;
    s->ms_scc = 0;
    if (nmarks < 2) return;
    (void)hipEventSynchronize(events[nmarks - 1]);
    for (size_t i = 0; i + 1 < nmarks; i++) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, events[i], events[i + 1]);
        if (mark_phase[i] >= 0 && phase_marks) s->ms_phase[mark_phase[i]] += ms;
        s->ms_scc += ms;
    }
}

uint32_t Engine::bpd(int W) const
{
    uint64_t b = ((uint64_t)g.V + 1) * (uint64_t)W / 256 + 1;
    return (uint32_t)(b > 2048 ? 2048 : b);
}

void Engine::init_vis(bool elect_first, bool use_live_bits)
{
    k_init_vis(stream, g, st, elect_first, use_live_bits ? live_bits() : nullptr);
    do_seeded = elect_first;
}

// The level-L arguments of direction d (0 = FW, 1 = BW): queues alternate
// with the level, counts and hub lists rotate through three slots.
DirArgs Engine::dir_args(int d, uint32_t level) const
{
    DirArgs D;
    const int in = level & 1, out = in ^ 1;
    const int hub = d ? C_BIGB : C_BIGF;
    const int cur = level % 3, prev = (level + 2) % 3, next = (level + 1) % 3;
    D.R = d ? g.Br : g.Fr;
    D.C = d ? g.Bc : g.Fc;
    D.vis = d ? st.bvis : st.fvis;
    D.qin = d ? st.qb[in] : st.qf[in];
    D.qout = d ? st.qb[out] : st.qf[out];
    D.cnt = st.ctr + CT(d ? C_BWCNT : C_FWCNT);
    D.cslot = st.ctr + CSLOT(d ? C_BWCNT : C_FWCNT, 0);
    D.big = d ? st.bigb[cur] : st.bigf[cur];
    D.bigoff = d ? st.bigoff_b[cur] : st.bigoff_f[cur];
    D.bigc = reinterpret_cast<unsigned long long *>(st.ctr + CT(hub + cur));
    D.pbig = d ? st.bigb[prev] : st.bigf[prev];
    D.pbigoff = d ? st.bigoff_b[prev] : st.bigoff_f[prev];
    D.pbigc = reinterpret_cast<unsigned long long *>(st.ctr + CT(hub + prev));
    D.bigc_zero = reinterpret_cast<unsigned long long *>(st.ctr + CT(hub + next));
    return D;
}

// FW and/or BW search from the seeded queues until both frontiers are
// empty, `K` levels per look at the counters (reference mappings and LB).
scc_status Engine::reach(bool fw, bool bw, bool single, int expand, int W, uint32_t *levels_out)
{
    if (!fw && !bw) return SCC_OK;
    if (expand == SCC_EXPAND_DO) return reach_do(fw, bw, single, levels_out);
    do_seeded = false;   // the seeded state is for a direction-optimised search only
    // levels queued past the end of the search are no-ops, so a look at the
    // counters every 8 levels costs at most 7 empty launches
    const int K = check_every > 0 ? check_every : 8;
    zero_ctr(C_BIGF, 6);   // both directions' hub lists start empty
    const uint32_t blocks = bpd(expand == SCC_EXPAND_VWARP ? W : 1);
    uint32_t level = 0;
    for (;;) {
        LevelArgs a;
        std::memset(&a, 0, sizeof a);
        a.cls = st.cls;
        a.level = (int)level;
        if (fw) a.d[a.ndirs++] = dir_args(0, level);
        if (bw) a.d[a.ndirs++] = dir_args(1, level);
        k_expand_level(stream, a, expand, W, single, blocks);
        level++;
        if (level % K != 0) continue;
        scc_status s = read_ctr();
        if (s != SCC_OK) return s;
        bool any = false;
        for (int d = 0; d < 2; d++) {
            if (!(d ? bw : fw)) continue;
            const uint32_t nf = h_ctr[CT((d ? C_BWCNT : C_FWCNT) + level % 3)];
            const uint64_t hubs = *reinterpret_cast<const unsigned long long *>(
                h_ctr + CT((d ? C_BIGB : C_BIGF) + (level + 2) % 3)) >> 32;
            any |= nf || hubs;
        }
        if (!any) break;
        // every level claims at least one vertex, so V+1 levels bound a search
        if (level > limit()) return no_convergence("reachability");
    }
    if (levels_out) *levels_out += level;
    return SCC_OK;
}

// Direction-optimised search.  Each level is one launch whose workgroups
// choose top-down or bottom-up per direction on the device (dir_decide), so
// the host queues a window of levels per look at the counters: kDoWindow
// levels on the full grid (kFirstWindow for a single-class search, the
// first of a decomposition: power-law graphs finish it within that many
// levels), or kSpecLevels levels on a grid sized for a few times the
// current frontier while every direction is top-down, has no hubs pending,
// grew at most kSpecRise-fold since the previous look and would have to
// grow kSpecGrowth-fold more to reach the bottom-up switch point
// (high-diameter graphs run thousands of small levels; such a window may
// not switch to bottom-up).  From the second look on the windows are
// pipelined: the next one is queued before the host reads the counters of
// the last, so the GPU does not wait for the host (its parameters come from
// the look before); a search that ends in its first window pays nothing
// for that.
// Levels queued past the end of the search are no-ops.
static constexpr int kDoWindow = 4;
// 6: the first searches of R-MAT scales 22 and 24 take 6 levels, so a
// window of 8 ran two no-op launches (about 4.5 us each on 2,048
// workgroups); a first search of 7 or more levels pays one more look
static constexpr int kFirstWindow = 6;
static constexpr int kSpecLevels = 16;
static constexpr int64_t kSpecGrowth = 8;
static constexpr int64_t kSpecRise = 2;

scc_status Engine::reach_do(bool fw, bool bw, bool single, uint32_t *levels_out)
{
    int dirs[2], nd = 0;   // argument position -> direction
    if (fw) dirs[nd++] = 0;
    if (bw) dirs[nd++] = 1;
    const uint32_t full = do_blocks(g.V, nd);
    if (!(do_seeded && nd == 2)) k_do_begin(stream, st.ctr, nd);
    do_seeded = false;
    DoArgs a;
    std::memset(&a, 0, sizeof a);
    a.cls = st.cls;
    a.ctr = st.ctr;
    a.V = g.V;
    a.ndirs = nd;
    a.qcap = slot_cap(g.V, full);   // bottom-up levels run on the full grid only
    a.count = search_stats && single ? 1 : 0;   // the first search of a run with sizes
    for (int i = 0; i < nd; i++) {
        a.PR[i] = dirs[i] ? g.Fr : g.Br;
        a.PC[i] = dirs[i] ? g.Fc : g.Bc;
        a.PF[i] = dirs[i] ? g.F1 : g.B1;
    }
    uint32_t level = 0;
    auto queue = [&](int window, uint32_t blocks) {
        a.bu_ok = blocks == full;
        for (int k = 0; k < window; k++, level++) {
            for (int i = 0; i < nd; i++) a.d[i] = dir_args(dirs[i], level);
            a.level = (int)level;
            k_do_level(stream, a, single, blocks);
        }
    };
    queue(single ? kFirstWindow : kDoWindow, full);
    uint32_t seen = level, seq = publish_ctr();
    int window = 0;   // nothing is queued ahead of the first look
    bool narrow_ok = true;
    uint32_t blocks = full;
    int64_t nf_prev[2] = {0, 0};
    for (;;) {
        const bool ahead = window > 0;
        uint32_t seen_next = seen, seq_next = seq;
        if (ahead) {
            queue(window, blocks);
            seen_next = level;
            seq_next = publish_ctr();
        }
        scc_status s = wait_ctr(seq);
        if (s != SCC_OK) return s;
        // the state level `seen` would compute: levels 0..seen-1 ran
        const int64_t nlive = h_ctr[CT(C_NLIVE)];
        bool any = false, spec = true;
        int64_t nf_max = 0;
        uint32_t end = 0;
        for (int i = 0; i < nd; i++) {
            const int d = dirs[i];
            const uint32_t last = seen - 1;
            const DirState prev = load_dir_state(h_ctr, last == 0 ? 0 : last & 1, i);
            const uint32_t nf = h_ctr[CT((d ? C_BWCNT : C_FWCNT) + seen % 3)];
            const uint32_t hubs = (uint32_t)(*reinterpret_cast<const unsigned long long *>(
                                                 h_ctr + CT((d ? C_BIGB : C_BIGF) + last % 3)) >> 32);
            const DirState next = dir_decide(prev, nf, hubs, (uint32_t)nlive, seen, true);
            if (trace_levels)
                std::fprintf(stderr, "[scc reach] after level %u: %s %s %u (hubs %u), visited %llu\n", last,
                             d ? "bw" : "fw",
                             next.mode == DIR_OFF ? "done" : next.mode == DIR_BU ? "BU" : "TD",
                             prev.mode == DIR_OFF ? 0u : nf, prev.mode == DIR_TD ? hubs : 0u,
                             (unsigned long long)next.visited);
            if (next.mode == DIR_OFF) {
                end = std::max(end, next.end);
                continue;
            }
            any = true;
            const int64_t remaining = nlive - (int64_t)next.visited;
            if (next.mode == DIR_BU || (prev.mode == DIR_TD && hubs) || !nf_prev[i] ||
                (int64_t)nf > nf_prev[i] * kSpecRise || (int64_t)nf * kToBottomUp * kSpecGrowth > remaining)
                spec = false;
            nf_prev[i] = nf;
            nf_max = std::max<int64_t>(nf_max, nf);
        }
        if (!any) {
            // a window queued meanwhile runs as no-ops; its counters are
            // published into the other buffer, which nothing waits for
            if (levels_out) *levels_out += end + h_ctr[CT(C_NLEVELS)] + h_ctr[CT(C_NLEVELS) + 1];
            return SCC_OK;
        }
        // every level claims at least one vertex, so V+1 levels bound a search
        if (seen > limit()) return no_convergence("reachability");
        // Narrow frontiers: every active direction top-down from a top-down
        // level, no hubs pending, a frontier that fits the workgroup's LDS.
        // The narrow kernel continues from `level` (everything before it is
        // queued) and rewrites that level's input; nothing is queued behind
        // it until the host has read where it stopped.  After a narrow run at
        // least one window of device-wide levels runs before the next one.
        // A single-class search does not use it: its top-down levels go on
        // expanding narrow frontiers inside the device-wide launch (local
        // continuation), which is faster on every high-diameter graph
        // measured (a 20,000-vertex cycle 50 -> 25 ms, the 4096^2 grid
        // 29 -> 27.6 ms) and leaves power-law searches unchanged.
        bool narrow = narrow_ok && !single;
        for (int i = 0; i < nd && narrow; i++) {
            const int d = dirs[i];
            const uint32_t last = seen - 1;
            const DirState prev = load_dir_state(h_ctr, last == 0 ? 0 : last & 1, i);
            if (prev.mode == DIR_OFF) continue;
            const uint32_t nf = h_ctr[CT((d ? C_BWCNT : C_FWCNT) + seen % 3)];
            const uint32_t hubs = (uint32_t)(*reinterpret_cast<const unsigned long long *>(
                                                 h_ctr + CT((d ? C_BIGB : C_BIGF) + last % 3)) >> 32);
            narrow = prev.mode == DIR_TD && !hubs && nf <= narrow_cap();
        }
        narrow_ok = true;
        if (narrow) {
            for (int i = 0; i < nd; i++) a.d[i] = dir_args(dirs[i], level);
            a.level = (int)level;
            k_narrow_levels(stream, a, single);
            seen = level;
            seq = publish_ctr();
            window = 0;
            narrow_ok = false;
            nf_prev[0] = nf_prev[1] = 0;
            continue;
        }
        window = spec ? kSpecLevels : kDoWindow;
        blocks = spec ? (uint32_t)std::min<int64_t>(full, std::max<int64_t>(32, (nf_max * 4 + 255) / 256)) : full;
        if (!ahead) {   // after the first look: queue now, ahead from then on
            queue(window, blocks);
            seen_next = level;
            seq_next = publish_ctr();
        }
        seen = seen_next;
        seq = seq_next;
    }
}

// trim-1 to its fixpoint (scc_WCC.cu:680-687), or one pass.  After the
// first sweep each pass re-checks the neighbours of the previous pass's
// removals (listed alternately in the FW queue buffers, which are free
// outside a search) until a pass removes nothing; a pass that follows one
// removing more than 1/kTrimSweepShare of the vertices sweeps instead.  The
// kernel makes that choice from the list count, so the host queues
// kTrimQueued passes per look at the counters.
static constexpr uint32_t kTrimSweepShare = 32;
static constexpr uint32_t kTrimQueued = 4;

scc_status Engine::trim1(bool fixpoint, bool single, bool none_dead, bool wait, bool labels)
{
    mark(SCC_PHASE_TRIM1);
    // a fresh state comes straight from the state initialisation, which
    // zeroed every counter line (the removal-list counts and their slots too)
    if (!none_dead) zero_ctr(C_TLIST, 3);
    // the single-class live bitmap lives in the byte scratch array, free
    // during a decomposition; the first pivot round reads it
    uint32_t *live = single ? live_bits() : nullptr;
    const bool degree_first = fixpoint && single && none_dead;
    const bool eager = labels || !degree_first;   // lazy: the live bitmap alone (no class or label writes)
    k_trim1(stream, g, st, st.qf[1], st.qf[0], 0u, 0, kTrimSweepShare, live, degree_first, eager);
    uint32_t pass = 1;
    if (!wait) {
        // one batch of passes and no look at the counters: the SCCs do not
        // depend on how far trimming got (what a fixpoint would still remove,
        // the phases after the first search remove), and the GPU goes on to
        // the next phase without waiting for the host
        for (uint32_t k = 0; fixpoint && k < kTrimQueued; k++, pass++) {
            if (++trim_epoch == 0) {
                k_fill_u32(stream, st.stamp, 0u, (uint64_t)g.V + 1);
                trim_epoch = 1;
            }
            k_trim1(stream, g, st, st.qf[(pass + 1) & 1], st.qf[pass & 1], trim_epoch, pass, kTrimSweepShare, live,
                    true, eager);
        }
        return SCC_OK;
    }
    while (fixpoint) {
        for (uint32_t k = 0; k < kTrimQueued; k++, pass++) {
            if (++trim_epoch == 0) {   // wrapped: no stale stamp may match a new epoch
                k_fill_u32(stream, st.stamp, 0u, (uint64_t)g.V + 1);
                trim_epoch = 1;
            }
            k_trim1(stream, g, st, st.qf[(pass + 1) & 1], st.qf[pass & 1], trim_epoch, pass, kTrimSweepShare, live,
                    true, eager);
        }
        scc_status s = read_ctr();
        if (s != SCC_OK) return s;
        // the next pass's input: what the last queued pass removed
        if (!h_ctr[CT(C_TLIST + pass % 3)]) break;
        // every pass but the last removes a vertex
        if (pass > limit()) return no_convergence("trim-1");
    }
    if (!fixpoint) return read_ctr();
    return SCC_OK;
}

uint64_t Engine::limit() const { return (uint64_t)g.V + 8; }

scc_status Engine::no_convergence(const char *what)
{
    return fail(SCC_ERR_DEVICE, "%s did not converge within %llu iterations", what,
                (unsigned long long)limit());
}

// One pivot round: elect, seed, FW||BW, update.  Returns live count.
scc_status Engine::pivot_round(bool first, int expand, int W, scc_stats *stats, uint32_t *live, int canon,
                               bool use_live_bits, bool lazy)
{
    lazy = lazy && first;
    mark(SCC_PHASE_PIVOT);
    init_vis(first, first && use_live_bits);
    if (!first) {
        zero_ctr(C_NPIV, 1);
        k_elect(stream, g, st);
        k_select(stream, g, st, 1);
    }
    mark(SCC_PHASE_REACH, first);   // the first search is always timed
    const int first_mark = (int)nmarks - 1;
    scc_status s = reach(true, true, first, expand, W, &stats->levels);
    if (s != SCC_OK) return s;
    mark(SCC_PHASE_UPDATE, first);
    if (first) {
        first_reach_marks[0] = first_mark;
        first_reach_marks[1] = (int)nmarks - 1;
        if (search_stats) k_reach_stats(stream, g, st);
    }
    // the first round of a decomposition follows the state initialisation,
    // which zeroed C_LIVE, and has one pivot at most
    if (!first) zero_ctr(C_LIVE, 1);
    // the first round's update also takes its SCC's minimum (one pivot)
    if (lazy && use_live_bits) k_update_words(stream, g, st, live_bits());
    else k_update(stream, g, st, first, first && use_live_bits ? live_bits() : nullptr, lazy);
    if (!first) k_clear_piv(stream, st, g.V + 1);   // the first round's update clears its pivot's entries
    if (canon >= 0) {
        // the labels of a decomposition this round finished, in the same
        // look at the counters (a no-op on the device if a vertex is live);
        // in the first round the only non-final label is the pivot's, whose
        // minimum the update left in aux
        mark(SCC_PHASE_LABEL);
        if (lazy && canon == 0) {
            k_label_from_bits(stream, g, st);
        } else if (first) {
            k_canon_apply(stream, g, st, true);
            if (canon) k_sizes(stream, g, st, true);
        } else {
            k_canon(stream, g, st, canon != 0, true);
        }
        mark(-1);
    }
    if ((s = read_ctr()) != SCC_OK) return s;
    *live = h_ctr[CT(C_LIVE)];
    if (first) stats->first_pivot = h_ctr[CT(C_PIV0)];
    return SCC_OK;
}

scc_status Engine::min_label()
{
    mark(SCC_PHASE_WCC);
    k_minlabel_init(stream, g, st);
    scc_status s;
    uint64_t it = 0;
    do {
        zero_ctr(C_CHANGED, 1);
        k_minlabel_prop(stream, g, st);
        if ((s = read_ctr()) != SCC_OK) return s;
        if (++it > limit()) return no_convergence("min-label propagation");
    } while (h_ctr[CT(C_CHANGED)]);
    k_minlabel_finish(stream, g, st);
    return SCC_OK;
}

scc_status Engine::color_loop(int expand, int W, scc_stats *stats)
{
    scc_status s;
    for (;;) {
        if ((s = read_ctr()) != SCC_OK) return s;
        if (!h_ctr[CT(C_LIVE)]) break;
        if (++stats->rounds > limit()) return no_convergence("colouring rounds");
        mark(SCC_PHASE_COLOR);
        k_color_init(stream, g, st);
        uint64_t it = 0;
        do {
            zero_ctr(C_CHANGED, 1);
            k_color_prop(stream, g, st);
            if ((s = read_ctr()) != SCC_OK) return s;
            if (++it > limit()) return no_convergence("colour propagation");
        } while (h_ctr[CT(C_CHANGED)]);
        init_vis();
        zero_ctr(C_NPIV, 1);
        k_color_roots(stream, g, st);
        mark(SCC_PHASE_REACH);
        if ((s = reach(true, false, false, expand, W, &stats->levels)) != SCC_OK) return s;
        mark(SCC_PHASE_UPDATE);
        zero_ctr(C_LIVE, 1);
        k_color_finish(stream, g, st);
    }
    return SCC_OK;
}

scc_status Engine::run(const scc_opts &o, scc_stats *out)
{
    if (!has_graph) return fail(SCC_ERR_NOTREADY, "no graph uploaded");
    HIP_TRY(hipSetDevice(device));
    const char A = o.algo;
    if (A != 'g' && A != 'h' && A != 'x' && A != 'y' && A != 'd')
        return fail(SCC_ERR_INVALID, "unknown algorithm '%c'", A);
    if (A == 'd' && !has_part)
        return fail(SCC_ERR_NOTREADY, "detectSCC needs a partition vector (scc_partition_upload)");
    int W = o.warp;
    if (W != 1 && W != 2 && W != 4 && W != 8 && W != 16 && W != 32 && W != 64)
        return fail(SCC_ERR_INVALID, "warp size %d not in {1,2,4,8,16,32,64}", W);
    int expand = o.expand;
    // The automatic mode is the direction-optimised search for every
    // algorithm: the labels do not depend on the expansion, and on R-MAT
    // scale 22 it is 23-42x faster than the reference's mappings, which stay
    // selectable (SCC_EXPAND_VERTEX, or SCC_EXPAND_VWARP with W lanes).
    if (expand == SCC_EXPAND_AUTO) expand = SCC_EXPAND_DO;
    if (expand != SCC_EXPAND_VERTEX && expand != SCC_EXPAND_VWARP && expand != SCC_EXPAND_LB &&
        expand != SCC_EXPAND_DO)
        return fail(SCC_ERR_INVALID, "unknown expand mode %d", expand);
    // the reference uses the virtual warp only for the first pivot's search
    // (scc_WCC.cu:209-317); the main loop uses vertex-per-thread
    const int first_expand = expand;
    const int loop_expand = (expand == SCC_EXPAND_VWARP) ? SCC_EXPAND_VERTEX : expand;
    check_every = o.check_every;
    search_stats = o.sizes != 0;
    phase_marks = o.phase_times != 0;

    scc_stats stats;
    std::memset(&stats, 0, sizeof stats);
    nmarks = 0;
    first_reach_marks[0] = first_reach_marks[1] = -1;
    mark(-1);
    const bool hong = (A == 'g' || A == 'x');
    // the first trim-1 pass of Hong and detectSCC starts from the fresh
    // state and writes the per-vertex words itself
    const bool fresh_trim = o.trim1 && (hong || A == 'd');
    // The first-neighbour and degree-product arrays belong to the
    // decomposition, not to the upload: every run builds them from the CSR.
    // The fresh trim's degree pass builds them for the vertices it keeps;
    // any other path builds them for every vertex first.
    if (!fresh_trim) k_first_arcs(stream, g);
    k_init_state(stream, g, st, nullptr, !fresh_trim);   // also zeroes the counters
    scc_status s;
    uint32_t live = g.V;
    bool labelled = false;   // the first pivot round finished the decomposition and its labels
    if (A == 'd') {
        if ((s = detect_phases(o, loop_expand, W, &stats, false)) != SCC_OK) return s;
    } else if (hong) {
        // lazy labels: the degree pass and the first update write no labels
        // (nor the members' class words); a decomposition the first round
        // finishes takes its labels from the search's bitmaps, otherwise a
        // fix-up writes what the eager round would have (sizes needs them)
        const bool lazy = o.trim1 && !o.sizes;
        if (o.trim1 && (s = trim1(true, true, true, false, !lazy)) != SCC_OK) return s;
        // the single-class trim just left an exact live bitmap
        if ((s = pivot_round(true, first_expand, W, &stats, &live, o.sizes != 0 ? 1 : 0, o.trim1 != 0, lazy)) !=
            SCC_OK)
            return s;
        labelled = live == 0;
        if (live && lazy) k_label_fixup(stream, g, st, live_bits());
        // with no live vertex left (a power-law graph's first search often
        // takes everything trim-1 left) every later phase removes nothing,
        // elects nothing and changes no label
        if (live) {
            if (o.trim1 && (s = trim1(true)) != SCC_OK) return s;
            if (o.trim2) {
                mark(SCC_PHASE_TRIM2);
                k_trim2(stream, g, st);
            }
            if (o.trim1 && (s = trim1(true)) != SCC_OK) return s;
            if ((s = min_label()) != SCC_OK) return s;
            for (;;) {
                // a round with no live vertex elects nothing and reports live = 0
                if ((s = pivot_round(false, loop_expand, W, &stats, &live)) != SCC_OK) return s;
                if (!live) break;
                // each round finishes at least its pivots' SCCs
                if (++stats.rounds > limit()) return no_convergence("pivot rounds");
            }
        }
    } else {
        // vSlota/wSlota: one trim-1 pass, -q ignored (scc_Coloring.cu:507-508)
        if (o.trim1 && (s = trim1(false, true, false, false)) != SCC_OK) return s;
        if ((s = pivot_round(true, first_expand, W, &stats, &live, -1, o.trim1 != 0)) != SCC_OK) return s;
        if ((s = color_loop(loop_expand, W, &stats)) != SCC_OK) return s;
    }
    if (!labelled) {
        if (A != 'd') {
            mark(SCC_PHASE_LABEL);
            k_canon(stream, g, st, o.sizes != 0);
        } else if (o.sizes) {   // detectSCC labels come out of the merge already canonical
            mark(SCC_PHASE_LABEL);
            k_sizes(stream, g, st);
        }
        mark(-1);
        if ((s = read_ctr()) != SCC_OK) return s;
    }
    if (h_ctr[CT(C_ANOM)])
        return fail(SCC_ERR_DEVICE, "internal invariant violated %u times", h_ctr[CT(C_ANOM)]);
    collect_times(&stats);
    if (first_reach_marks[0] >= 0) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, events[first_reach_marks[0]], events[first_reach_marks[1]]);
        stats.ms_first_reach = ms;
    }
    auto u64 = [&](int i) { return (uint64_t)h_ctr[CT(i)] | ((uint64_t)h_ctr[CT(i) + 1] << 32); };
    stats.num_sccs = h_ctr[CT(C_NSCC)];
    stats.num_trim1 = h_ctr[CT(C_TRIM1)];
    stats.num_trim2 = h_ctr[CT(C_TRIM2)];
    stats.num_pivot = h_ctr[CT(C_PIVOT)];
    stats.largest = h_ctr[CT(C_LARGEST)];
    stats.first_fw_vertices = u64(C_FWV);
    stats.first_fw_arcs = u64(C_FWE);
    stats.first_bw_vertices = u64(C_BWV);
    stats.first_bw_arcs = u64(C_BWE);
    for (int k = 0; k < 8; k++) stats.first_work[k] = u64(C_WORK + k);
    ran = true;
    if (out) *out = stats;
    return SCC_OK;
}

// detectSCC (scc_WCC.cu:1000-1617): global trims (1122-1173), local SCCs
// per (partition, class) (1175-1269), then the merge as a quotient-graph
// SCC computed by a second engine on the same device.  With local_only the
// canonical local labels are left in lab.
scc_status Engine::detect_phases(const scc_opts &o, int expand, int W, scc_stats *stats, bool local_only)
{
    if (!has_part) return fail(SCC_ERR_NOTREADY, "detectSCC needs a partition vector");
    scc_status s;
    if (o.trim1 && (s = trim1(true, true, true)) != SCC_OK) return s;
    if (o.trim2) {
        // after the single-class trim-1 every live vertex is in class 0 and
        // its live bitmap is exact (nothing reads the bitmap after trim-2)
        mark(SCC_PHASE_TRIM2);
        k_trim2(stream, g, st, o.trim1 ? live_bits() : nullptr);
    }
    // every live vertex is still in class 0 (trim-2 only removes), so the
    // second fixpoint runs single-class too: its first pass rebuilds the
    // live bitmap from the class words, and it tests neighbours in the bitmap
    if (o.trim1 && (s = trim1(true, true)) != SCC_OK) return s;
    mark(SCC_PHASE_LOCAL);
    const size_t nv = (size_t)g.V + 1;
    uint32_t *tbits, *isrep, *qid, *rep, *cout, *cin, *Qr, *QBr, *tiles;
    if ((s = scratch(buf[0], st.nwords, &tbits)) != SCC_OK) return s;
    k_local_init(stream, g, st, tbits);
    uint32_t live = 0;
    for (;;) {
        if ((s = pivot_round(false, expand, W, stats, &live)) != SCC_OK) return s;
        if (!live) break;
        if (++stats->rounds > limit()) return no_convergence("local rounds");
    }
    mark(SCC_PHASE_LABEL);
    k_canon(stream, g, st, false);
    if (local_only) return SCC_OK;
    mark(SCC_PHASE_MERGE);
    if ((s = scratch(buf[1], nv, &isrep)) != SCC_OK || (s = scratch(buf[2], nv, &qid)) != SCC_OK ||
        (s = scratch(buf[3], nv, &rep)) != SCC_OK || (s = scratch(buf[4], nv + 1, &cout)) != SCC_OK ||
        (s = scratch(buf[5], nv + 1, &cin)) != SCC_OK || (s = scratch(buf[6], nv + 1, &Qr)) != SCC_OK ||
        (s = scratch(buf[7], nv + 1, &QBr)) != SCC_OK ||
        (s = scratch(buf[8], scan_tile_count(nv + 1), &tiles)) != SCC_OK)
        return s;
    k_q_mark(stream, g, st, tbits, isrep);
    k_exscan(stream, isrep, qid, nv, tiles);
    uint32_t last[2];
    HIP_TRY(hipMemcpyAsync(&last[0], qid + g.V, 4, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipMemcpyAsync(&last[1], isrep + g.V, 4, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    const uint32_t QV = last[0] + last[1];
    const uint32_t *qlab = nullptr;
    if (QV) {
        // quotient arcs as sorted 64-bit keys: each vertex's quotient id,
        // the keys of the arcs between different local SCCs, and two radix
        // sorts into the rows of both directions.  The keys go straight into
        // the (grow-only) buffer an earlier merge left; only a first merge,
        // or one whose keys did not fit, counts them first.
        uint32_t *qv;
        unsigned long long *qn;
        if ((s = scratch(buf[11], nv, &qv)) != SCC_OK || (s = scratch(buf[12], 1, &qn)) != SCC_OK) return s;
        k_q_vid(stream, g, st, tbits, qid, qv);
        auto read_count = [&](unsigned long long *out) -> scc_status {
            HIP_TRY(hipMemcpyAsync(out, qn, sizeof *out, hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipStreamSynchronize(stream));
            return SCC_OK;
        };
        unsigned long long qe = 0;
        uint64_t *keys = static_cast<uint64_t *>(buf[13].p);
        uint64_t kcap = buf[13].cap / sizeof(uint64_t);
        for (int attempt = 0;; attempt++) {
            k_fill_u32(stream, reinterpret_cast<uint32_t *>(qn), 0u, 2);
            if (kcap) k_q_keys(stream, true, g, qv, qn, keys, kcap);
            else k_q_keys(stream, false, g, qv, qn, nullptr, 0);
            if ((s = read_count(&qe)) != SCC_OK) return s;
            if (qe > 0xFFFFFFFFull) return fail(SCC_ERR_INVALID, "quotient graph has %llu arcs", qe);
            if (kcap && qe <= kcap) break;   // every key written
            if (attempt == 2) return fail(SCC_ERR_DEVICE, "quotient arc count changed between passes");
            if ((s = scratch(buf[13], std::max<uint64_t>(qe, 1), &keys)) != SCC_OK) return s;
            kcap = buf[13].cap / sizeof(uint64_t);
        }
        const uint32_t QE = (uint32_t)qe;
        if (QE) {
            uint32_t *Qc, *QBc;
            uint64_t *sorted;
            void *tmp;
            const size_t tmp_bytes = csr_sort_bytes(QV, QE);
            if ((s = scratch(buf[9], QE, &Qc)) != SCC_OK || (s = scratch(buf[10], QE, &QBc)) != SCC_OK ||
                (s = scratch(buf[14], QE, &sorted)) != SCC_OK ||
                (s = scratch(buf[15], tmp_bytes, reinterpret_cast<uint8_t **>(&tmp))) != SCC_OK)
                return s;
            if ((s = csr_from_keys(stream, QV, QE, keys, sorted, tmp, tmp_bytes, Qr, Qc, QBr, QBc)) != SCC_OK)
                return s;
            k_q_rep(stream, g, isrep, qid, rep);
            if (!sub) {
                sub.reset(new Engine());
                if ((s = sub->init(device)) != SCC_OK) return s;
            }
            if ((s = sub->upload_device(QV, QE, Qr, Qc, QBr, QBc, stream)) != SCC_OK) return s;
            scc_opts so;
            scc_opts_default(&so);
            so.algo = 'g';
            so.expand = SCC_EXPAND_DO;
            so.sizes = 0;
            scc_stats sst;
            if ((s = sub->run(so, &sst)) != SCC_OK) return s;
            qlab = sub->st.lab;
            stats->levels += sst.levels;
        }
    }
    zero_ctr(C_NSCC, 1);
    k_q_map(stream, g, st, tbits, qid, qlab, rep);
    return SCC_OK;
}

// The condensation DAG of the last run: node k = the SCC with the k-th
// smallest canonical label, an arc per distinct pair of SCCs joined by an
// arc.  The quotient kernels of the detectSCC merge count and place the arcs
// between different SCCs on the device (no vertex excluded: an empty
// finished bitmap); the host sorts, removes parallel arcs and builds the
// reverse rows.
scc_status Engine::condense(scc_host_graph *out, uint32_t *node_of)
{
    if (!ran) return fail(SCC_ERR_NOTREADY, "no completed run");
    std::memset(out, 0, sizeof *out);
    const size_t nv = (size_t)g.V + 1;
    uint32_t *none, *isrep, *qid, *cout, *cin, *Qr, *QBr, *tiles;
    scc_status s;
    if ((s = scratch(buf[0], st.nwords, &none)) != SCC_OK || (s = scratch(buf[1], nv, &isrep)) != SCC_OK ||
        (s = scratch(buf[2], nv, &qid)) != SCC_OK || (s = scratch(buf[4], nv + 1, &cout)) != SCC_OK ||
        (s = scratch(buf[5], nv + 1, &cin)) != SCC_OK || (s = scratch(buf[6], nv + 1, &Qr)) != SCC_OK ||
        (s = scratch(buf[7], nv + 1, &QBr)) != SCC_OK ||
        (s = scratch(buf[8], scan_tile_count(nv + 1), &tiles)) != SCC_OK)
        return s;
    k_fill_u32(stream, none, 0u, st.nwords);
    k_q_mark(stream, g, st, none, isrep);
    k_exscan(stream, isrep, qid, nv, tiles);
    std::vector<uint32_t> hq(nv), hlab(nv), hrep(nv);
    HIP_TRY(hipMemcpyAsync(hq.data(), qid, sizeof(uint32_t) * nv, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipMemcpyAsync(hrep.data(), isrep, sizeof(uint32_t) * nv, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipMemcpyAsync(hlab.data(), st.lab, sizeof(uint32_t) * nv, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    const uint32_t QV = hq[g.V] + hrep[g.V];
    if (node_of) {
        node_of[0] = 0;
        for (uint32_t v = 1; v <= g.V; v++) node_of[v] = hq[hlab[v]] + 1;
    }
    std::vector<uint32_t> src, dst;
    if (QV) {
        k_fill_u32(stream, cout, 0, (uint64_t)QV + 2);
        k_fill_u32(stream, cin, 0, (uint64_t)QV + 2);
        k_q_arcs(stream, 0, g, st, none, qid, cout, cin, nullptr, nullptr);
        k_exscan(stream, cout, Qr, (uint64_t)QV + 2, tiles);
        k_exscan(stream, cin, QBr, (uint64_t)QV + 2, tiles);
        std::vector<uint32_t> hr((size_t)QV + 2);
        HIP_TRY(hipMemcpyAsync(hr.data(), Qr, sizeof(uint32_t) * hr.size(), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        const uint32_t QE = hr[QV + 1];
        if (QE) {
            uint32_t *Qc, *QBc;
            if ((s = scratch(buf[9], QE, &Qc)) != SCC_OK || (s = scratch(buf[10], QE, &QBc)) != SCC_OK)
                return s;
            HIP_TRY(hipMemcpyAsync(cout, Qr, sizeof(uint32_t) * ((size_t)QV + 2), hipMemcpyDeviceToDevice, stream));
            HIP_TRY(hipMemcpyAsync(cin, QBr, sizeof(uint32_t) * ((size_t)QV + 2), hipMemcpyDeviceToDevice, stream));
            k_q_arcs(stream, 1, g, st, none, qid, cout, cin, Qc, QBc);
            dst.resize(QE);
            HIP_TRY(hipMemcpyAsync(dst.data(), Qc, sizeof(uint32_t) * QE, hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipStreamSynchronize(stream));
            src.resize(QE);
            for (uint32_t q = 1; q <= QV; q++)
                for (uint32_t i = hr[q]; i < hr[q + 1]; i++) src[i] = q;
        }
    }
    scc_host_graph dag;
    const uint32_t none_arc = 0;
    if ((s = scc_csr_from_arcs(QV, src.size(), src.empty() ? &none_arc : src.data(),
                               dst.empty() ? &none_arc : dst.data(), &dag)) != SCC_OK)
        return s;
    // parallel arcs out: rows are sorted, so keep the first of each run
    auto dedupe = [&](uint32_t *R, uint32_t *C) {
        uint32_t w = 0;
        for (uint32_t v = 1; v <= QV; v++) {
            const uint32_t b = R[v], e = R[v + 1];
            R[v] = w;
            for (uint32_t i = b; i < e; i++)
                if (i == b || C[i] != C[i - 1]) C[w++] = C[i];
        }
        R[QV + 1] = w;
        return w;
    };
    const uint32_t E1 = dedupe(dag.Fr, dag.Fc), E2 = dedupe(dag.Br, dag.Bc);
    if (E1 != E2) {
        scc_host_graph_free(&dag);
        return fail(SCC_ERR_DEVICE, "condensation rows disagree (%u forward, %u reverse arcs)", E1, E2);
    }
    dag.E = E1;
    *out = dag;
    return SCC_OK;
}

scc_status Engine::labels(uint32_t *lab)
{
    if (!ran) return fail(SCC_ERR_NOTREADY, "no completed run");
    HIP_TRY(hipMemcpyAsync(lab, st.lab, sizeof(uint32_t) * ((size_t)g.V + 1), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return SCC_OK;
}

// The size histogram of the last run, on the device: sizes from the
// canonical labels (k_sizes), then the per-size counts and their non-zero
// entries (at most floor(sqrt(2V)) + 1 distinct sizes: the sizes of
// distinct SCCs sum to at most V).  *ms is the device time of the three
// passes plus the compaction.
scc_status Engine::histogram(uint64_t *pairs, uint32_t cap, uint32_t *npairs, double *ms)
{
    if (!ran) return fail(SCC_ERR_NOTREADY, "no completed run");
    HIP_TRY(hipSetDevice(device));
    uint32_t kmax = 1;
    while ((uint64_t)kmax * (kmax + 1) / 2 <= g.V) kmax++;
    uint64_t *dpairs;
    scc_status s;
    if ((s = scratch(buf[11], 2 * (size_t)kmax + 2, &dpairs)) != SCC_OK) return s;
    uint32_t *dn = reinterpret_cast<uint32_t *>(dpairs + 2 * (size_t)kmax);
    if (events.size() < 2) {
        for (size_t i = events.size(); i < 2; i++) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            events.push_back(e);
        }
    }
    nmarks = 0;   // the run's phase marks are already collected
    HIP_TRY(hipEventRecord(events[0], stream));
    k_fill_u32(stream, dn, 0u, 1);
    k_sizes(stream, g, st);
    k_size_histogram(stream, g, st, st.aux, dpairs, kmax, dn);
    HIP_TRY(hipEventRecord(events[1], stream));
    uint32_t n = 0;
    HIP_TRY(hipMemcpyAsync(&n, dn, sizeof n, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (n > kmax) return fail(SCC_ERR_DEVICE, "%u distinct SCC sizes exceed the bound %u", n, kmax);
    std::vector<uint64_t> h(2 * (size_t)n);
    if (n) HIP_TRY(hipMemcpy(h.data(), dpairs, sizeof(uint64_t) * h.size(), hipMemcpyDeviceToHost));
    float t = 0;
    (void)hipEventElapsedTime(&t, events[0], events[1]);
    if (ms) *ms = t;
    std::vector<std::pair<uint64_t, uint64_t>> v(n);
    for (uint32_t i = 0; i < n; i++) v[i] = {h[2 * i], h[2 * i + 1]};
    std::sort(v.begin(), v.end());
    for (uint32_t i = 0; i < n && i < cap; i++) {
        pairs[2 * i] = v[i].first;
        pairs[2 * i + 1] = v[i].second;
    }
    *npairs = n;
    return SCC_OK;
}

// --------------------------------------------------- phase entries ---

scc_status Engine::phase_begin(const uint8_t *dead)
{
    if (!has_graph) return fail(SCC_ERR_NOTREADY, "no graph uploaded");
    HIP_TRY(hipSetDevice(device));
    ran = false;
    phase_marks = true;   // the phase entry points time their phase
    if (dead) HIP_TRY(hipMemcpyAsync(d_tmp8, dead, (size_t)g.V + 1, hipMemcpyHostToDevice, stream));
    k_first_arcs(stream, g);
    k_init_state(stream, g, st, dead ? d_tmp8 : nullptr);   // also zeroes the counters
    return SCC_OK;
}

scc_status Engine::fetch8(uint8_t *out)
{
    HIP_TRY(hipMemcpyAsync(out, d_tmp8, (size_t)g.V + 1, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return SCC_OK;
}

scc_status Engine::fetch32(uint32_t *out, const uint32_t *src)
{
    HIP_TRY(hipMemcpyAsync(out, src, sizeof(uint32_t) * ((size_t)g.V + 1), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return SCC_OK;
}

scc_status Engine::phase_local(uint32_t *out)
{
    if (!has_part) return fail(SCC_ERR_NOTREADY, "no partition uploaded");
    scc_status s;
    if ((s = phase_begin(nullptr)) != SCC_OK) return s;
    nmarks = 0;
    scc_opts o;
    scc_opts_default(&o);
    // trim-2 would finish a 2-cycle whose ends lie in different partitions as
    // one pair: a global SCC but two local ones.  Trim-1 removes only
    // vertices that are singletons locally as well.
    o.trim2 = 0;
    scc_stats tmp;
    std::memset(&tmp, 0, sizeof tmp);
    if ((s = detect_phases(o, SCC_EXPAND_DO, 1, &tmp, true)) != SCC_OK) return s;
    if ((s = fetch32(out, st.lab)) != SCC_OK) return s;
    out[0] = 0;
    return SCC_OK;
}

scc_status Engine::phase_trim1(const uint8_t *dead, uint8_t *removed)
{
    scc_status s;
    if ((s = phase_begin(dead)) != SCC_OK) return s;
    nmarks = 0;
    if ((s = trim1(true, true, dead == nullptr)) != SCC_OK) return s;
    k_export_removed(stream, g, st, d_tmp8);
    if ((s = fetch8(removed)) != SCC_OK) return s;
    if (dead)
        for (uint32_t v = 0; v <= g.V; v++) removed[v] = removed[v] && !dead[v];
    return SCC_OK;
}

scc_status Engine::phase_trim2(const uint8_t *dead, uint32_t *pair)
{
    scc_status s;
    if ((s = phase_begin(dead)) != SCC_OK) return s;
    k_trim2(stream, g, st);
    std::vector<uint32_t> cls((size_t)g.V + 1);
    if ((s = fetch32(cls.data(), st.cls)) != SCC_OK) return s;
    if ((s = fetch32(pair, st.lab)) != SCC_OK) return s;
    for (uint32_t v = 0; v <= g.V; v++)
        pair[v] = (v == 0 || cls[v] != DEAD || (dead && dead[v])) ? 0u : (pair[v] & ~LAB_FINAL);
    return SCC_OK;
}

scc_status Engine::phase_first_pivot(const uint8_t *dead, uint32_t *pivot)
{
    scc_status s;
    if ((s = phase_begin(dead)) != SCC_OK) return s;
    init_vis(true);
    if ((s = read_ctr()) != SCC_OK) return s;
    *pivot = 0;
    if (h_ctr[CT(C_NPIV)] > 1) return fail(SCC_ERR_DEVICE, "first-pivot election chose %u vertices", h_ctr[CT(C_NPIV)]);
    if (h_ctr[CT(C_NPIV)] == 1) {
        HIP_TRY(hipMemcpyAsync(pivot, st.pivlist, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
    }
    return SCC_OK;
}

scc_status Engine::phase_reach(uint32_t src, int dir, int expand, int W, const uint8_t *dead,
                               uint8_t *out, uint32_t *levels, double *ms)
{
    if (!has_graph) return fail(SCC_ERR_NOTREADY, "no graph uploaded");
    if (src < 1 || src > g.V) return fail(SCC_ERR_INVALID, "source %u outside [1, %u]", src, g.V);
    if (dead && dead[src]) return fail(SCC_ERR_INVALID, "source %u is marked dead", src);
    if (expand == SCC_EXPAND_AUTO) expand = SCC_EXPAND_VERTEX;
    if (expand < SCC_EXPAND_VERTEX || expand > SCC_EXPAND_DO)
        return fail(SCC_ERR_INVALID, "unknown expand mode %d", expand);
    if (W != 1 && W != 2 && W != 4 && W != 8 && W != 16 && W != 32 && W != 64)
        return fail(SCC_ERR_INVALID, "warp size %d not in {1,2,4,8,16,32,64}", W);
    scc_status s;
    if ((s = phase_begin(dead)) != SCC_OK) return s;
    init_vis();
    k_seed_single(stream, st, src, dir);
    nmarks = 0;
    mark(SCC_PHASE_REACH);
    uint32_t lv = 0;
    if ((s = reach(dir == 0, dir != 0, true, expand, W, &lv)) != SCC_OK) return s;
    mark(-1);
    scc_stats tmp;
    collect_times(&tmp);
    if (ms) *ms = tmp.ms_scc;
    if (levels) *levels = lv;
    k_export_visited(stream, g, st, dir, d_tmp8);
    return fetch8(out);
}

scc_status Engine::phase_min_label(const uint8_t *dead, uint32_t *out)
{
    scc_status s;
    if ((s = phase_begin(dead)) != SCC_OK) return s;
    k_minlabel_init(stream, g, st);
    uint64_t it = 0;
    do {
        zero_ctr(C_CHANGED, 1);
        k_minlabel_prop(stream, g, st);
        if ((s = read_ctr()) != SCC_OK) return s;
        if (++it > limit()) return no_convergence("min-label propagation");
    } while (h_ctr[CT(C_CHANGED)]);
    if ((s = fetch32(out, st.cls)) != SCC_OK) return s;
    for (uint32_t v = 0; v <= g.V; v++)
        if (out[v] == DEAD || v == 0) out[v] = 0;
    return SCC_OK;
}

scc_status Engine::phase_max_color(const uint8_t *dead, uint32_t *out)
{
    scc_status s;
    if ((s = phase_begin(dead)) != SCC_OK) return s;
    k_color_init(stream, g, st);
    uint64_t it = 0;
    do {
        zero_ctr(C_CHANGED, 1);
        k_color_prop(stream, g, st);
        if ((s = read_ctr()) != SCC_OK) return s;
        if (++it > limit()) return no_convergence("colour propagation");
    } while (h_ctr[CT(C_CHANGED)]);
    return fetch32(out, st.aux);
}

} // namespace sccamd
