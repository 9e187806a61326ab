// batch_engine.hip -- host side of the batched form: device-resident batch
// layout, the persistent launches, the one-graph-per-launch baseline, the
// per-bucket autotuner and the C ABI for batches.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "batch.h"
#include "batch_engine.h"
#include "engine.h"
#include "host_util.h"

namespace sccamd {

#define BHIP(call)                                                                          \
    do {                                                                                    \
        hipError_t _e = (call);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail(SCC_ERR_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), \
                        __FILE__, __LINE__);                                                \
    } while (0)

BatchEngine::~BatchEngine() { release(); }

void BatchEngine::release()
{
    for (void *p : allocs)
        if (p) (void)hipFree(p);
    allocs.clear();
    if (stream2) (void)hipStreamDestroy(stream2);
    stream2 = nullptr;
    d = BatchDev();
    G = 0;
    capS = capA = 0;
    capG = 0;
}

template <class T> static hipError_t balloc(std::vector<void *> &allocs, T **p, size_t n)
{
    hipError_t e = hipMalloc((void **)p, sizeof(T) * (n ? n : 1));
    if (e == hipSuccess) allocs.push_back(*p);
    return e;
}

// Size/degree bucket of a graph (4 size classes x 3 degree classes).
static int bucket_of(uint32_t V, uint32_t E)
{
    int s = V < (1u << 10) ? 0 : V < (1u << 12) ? 1 : V < (1u << 14) ? 2 : 3;
    double deg = V ? (double)E / V : 0;
    int d = deg < 6 ? 0 : deg < 11 ? 1 : 2;
    return s * 3 + d;
}

// Batch layout and device buffers for graphs with V[i] vertices and E[i]
// arcs, rows of graph i at row_off[i] and columns at col_off[i] in the
// source; the rows go to 64-aligned slots voff[i].  Buffers grow only.
scc_status BatchEngine::prepare(int device, hipStream_t stream, uint32_t ng, const uint32_t *Vs, const uint32_t *Es,
                                const uint64_t *roff, const uint64_t *coff)
{
    dev = device;
    s1 = stream;
    BHIP(hipSetDevice(device));
    if (!stream2) BHIP(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
    ran = false;
    G = ng;
    V.assign(Vs, Vs + G);
    E.assign(Es, Es + G);
    row_off.assign(roff, roff + G + 1);
    voff.assign(G + 1, 0);
    eoff.assign(coff, coff + G + 1);
    nverts = narcs = 0;
    for (uint32_t i = 0; i < G; i++) {
        if (V[i] >= (1u << 30)) return fail(SCC_ERR_INVALID, "graph %u: V too large", i);
        if (eoff[i + 1] - eoff[i] != E[i] || row_off[i + 1] - row_off[i] < (uint64_t)V[i] + 2)
            return fail(SCC_ERR_INVALID, "graph %u: offsets do not match V/E", i);
        voff[i + 1] = voff[i] + (((uint64_t)V[i] + 2 + 63) / 64) * 64;
        nverts += V[i];
        narcs += E[i];
    }
    const uint64_t S = voff[G];
    const uint64_t A = eoff[G];
    // grow-only device buffers: a serving slot reuses them batch after batch
    if (S > capS || A > capA || G > capG) {
        for (void *p : allocs)
            if (p) (void)hipFree(p);
        allocs.clear();
        capS = std::max<uint64_t>(S, capS + capS / 4);
        capA = std::max<uint64_t>(A, capA + capA / 4);
        capG = std::max<uint32_t>(G, capG + capG / 4);
        hipError_t e = hipSuccess;
        auto chk = [&](hipError_t r) {
            if (r != hipSuccess && e == hipSuccess) e = r;
        };
        chk(balloc(allocs, &m_voff, capG + 1));
        chk(balloc(allocs, &m_eoff, capG + 1));
        chk(balloc(allocs, &m_V, capG));
        chk(balloc(allocs, &m_Fr, capS));
        chk(balloc(allocs, &m_Br, capS));
        chk(balloc(allocs, &m_Fc, capA));
        chk(balloc(allocs, &m_Bc, capA));
        chk(balloc(allocs, &d.cls, capS));
        chk(balloc(allocs, &d.lab, capS));
        chk(balloc(allocs, &d.aux, capS));
        chk(balloc(allocs, &d.fvis, capS / 32 + 2));
        chk(balloc(allocs, &d.bvis, capS / 32 + 2));
        chk(balloc(allocs, &d.piv, 3 * capS));
        chk(balloc(allocs, &d.q, 4 * capS));
        chk(balloc(allocs, &d.pivlist, capS));
        chk(balloc(allocs, &m_cfg, capG));
        chk(balloc(allocs, &m_order, capG));
        chk(balloc(allocs, &d.nscc, capG));
        chk(balloc(allocs, &d.tstart, capG));
        chk(balloc(allocs, &d.tend, capG));
        chk(balloc(allocs, &d.next, 64));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            release();
            return fail(SCC_ERR_NOMEM, "Error: Not enough memory on GPU for the batch (%llu slots, %llu arcs)",
                        (unsigned long long)S, (unsigned long long)A);
        }
    }
    return SCC_OK;
}

// Metadata to the device and the per-graph tables of the launches.
scc_status BatchEngine::finish_upload()
{
    BHIP(hipMemcpyAsync(m_voff, voff.data(), sizeof(uint64_t) * (G + 1), hipMemcpyHostToDevice, s1));
    BHIP(hipMemcpyAsync(m_eoff, eoff.data(), sizeof(uint64_t) * (G + 1), hipMemcpyHostToDevice, s1));
    BHIP(hipMemcpyAsync(m_V, V.data(), sizeof(uint32_t) * G, hipMemcpyHostToDevice, s1));
    BHIP(hipStreamSynchronize(s1));
    d.G = G;
    d.voff = m_voff;
    d.eoff = m_eoff;
    d.Vn = m_V;
    d.Fr = m_Fr;
    d.Br = m_Br;
    d.Fc = m_Fc;
    d.Bc = m_Bc;
    d.cfg = m_cfg;
    d.order = m_order;
    bucket.resize(G);
    for (uint32_t i = 0; i < G; i++) bucket[i] = bucket_of(V[i], E[i]);
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || rate_khz <= 0)
        rate_khz = 100000;
    ticks_per_ms = (double)rate_khz;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    num_cus = cus;
    tuned = false;
    return SCC_OK;
}

scc_status BatchEngine::upload(int device, hipStream_t stream, const scc_batch_host *b)
{
    if (!b || !b->num_graphs || !b->V || !b->E || !b->row_off || !b->col_off || !b->Fr || !b->Br ||
        !b->Fc || !b->Bc)
        return fail(SCC_ERR_INVALID, "incomplete batch description");
    // each graph's rows must be a valid reference-layout CSR
    for (uint32_t i = 0; i < b->num_graphs; i++) {
        const uint32_t *Fr = b->Fr + b->row_off[i], *Br = b->Br + b->row_off[i];
        if (b->row_off[i + 1] - b->row_off[i] < (uint64_t)b->V[i] + 2) break;   // prepare() reports it
        if (Fr[0] || Fr[1] || Br[0] || Br[1] || Fr[b->V[i] + 1] != b->E[i] || Br[b->V[i] + 1] != b->E[i])
            return fail(SCC_ERR_INVALID, "graph %u: row arrays are not in the reference layout", i);
    }
    scc_status st = prepare(device, stream, b->num_graphs, b->V, b->E, b->row_off, b->col_off);
    if (st != SCC_OK) return st;
    std::atomic<int> bad{0};
    parallel_chunks(G, [&](uint64_t lo, uint64_t hi, int) {
        for (uint64_t i = lo; i < hi; i++) {
            const uint32_t *Fr = b->Fr + row_off[i], *Br = b->Br + row_off[i];
            const uint32_t *Fc = b->Fc + eoff[i], *Bc = b->Bc + eoff[i];
            for (uint32_t v = 0; v <= V[i]; v++)
                if (Fr[v] > Fr[v + 1] || Br[v] > Br[v + 1]) bad = 1;
            for (uint32_t k = 0; k < E[i]; k++)
                if (Fc[k] < 1 || Fc[k] > V[i] || Bc[k] < 1 || Bc[k] > V[i]) bad = 1;
        }
    });
    if (bad) return fail(SCC_ERR_INVALID, "a batch graph has an invalid CSR");
    const uint64_t S = voff[G], A = eoff[G];
    // rows into their 64-aligned slots, columns as given
    stage.assign(2 * S, 0);
    for (int which = 0; which < 2; which++) {
        const uint32_t *src = which ? b->Br : b->Fr;
        uint32_t *dst = stage.data() + which * S;
        parallel_chunks(G, [&](uint64_t lo, uint64_t hi, int) {
            for (uint64_t i = lo; i < hi; i++)
                std::memcpy(dst + voff[i], src + row_off[i], sizeof(uint32_t) * ((size_t)V[i] + 2));
        });
    }
    BHIP(hipMemcpyAsync(m_Fr, stage.data(), sizeof(uint32_t) * S, hipMemcpyHostToDevice, s1));
    BHIP(hipMemcpyAsync(m_Br, stage.data() + S, sizeof(uint32_t) * S, hipMemcpyHostToDevice, s1));
    BHIP(hipMemcpyAsync(m_Fc, b->Fc, sizeof(uint32_t) * A, hipMemcpyHostToDevice, s1));
    BHIP(hipMemcpyAsync(m_Bc, b->Bc, sizeof(uint32_t) * A, hipMemcpyHostToDevice, s1));
    return finish_upload();
}

uint32_t BatchEngine::info(uint32_t *Vs, uint32_t *Es, uint64_t *roff, uint64_t *coff) const
{
    for (uint32_t i = 0; i < G; i++) {
        if (Vs) Vs[i] = V[i];
        if (Es) Es[i] = E[i];
    }
    for (uint32_t i = 0; i <= G && G; i++) {
        if (roff) roff[i] = row_off[i];
        if (coff) coff[i] = eoff[i];
    }
    return G;
}

// Rows of graph g (V+2 entries at src + roff[g]) into slot voff[g]: one
// workgroup per graph.
__global__ void batch_rows_kernel(const uint32_t *src, const uint64_t *roff, const uint64_t *dvoff,
                                  const uint32_t *Vn, uint32_t G, uint32_t *dst)
{
    for (uint32_t g = blockIdx.x; g < G; g += gridDim.x) {
        const uint32_t n = Vn[g] + 2;
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) dst[dvoff[g] + i] = src[roff[g] + i];
    }
}

// A batch already in device memory in the host layout (rows of graph i at
// row_off[i], V[i] + 2 entries each; columns at col_off[i]), valid by
// construction (the device generators).
scc_status BatchEngine::upload_device(int device, hipStream_t stream, uint32_t ng, const uint32_t *Vs,
                                      const uint32_t *Es, const uint64_t *roff, const uint64_t *coff,
                                      const uint64_t *d_roff, const uint32_t *dFr, const uint32_t *dBr,
                                      const uint32_t *dFc, const uint32_t *dBc)
{
    scc_status st = prepare(device, stream, ng, Vs, Es, roff, coff);
    if (st != SCC_OK) return st;
    BHIP(hipMemcpyAsync(m_voff, voff.data(), sizeof(uint64_t) * (G + 1), hipMemcpyHostToDevice, s1));
    BHIP(hipMemcpyAsync(m_V, V.data(), sizeof(uint32_t) * G, hipMemcpyHostToDevice, s1));
    const uint32_t blocks = std::min<uint32_t>(G, 8192);
    batch_rows_kernel<<<blocks, 256, 0, s1>>>(dFr, d_roff, m_voff, m_V, G, m_Fr);
    batch_rows_kernel<<<blocks, 256, 0, s1>>>(dBr, d_roff, m_voff, m_V, G, m_Br);
    if (eoff[G]) {
        BHIP(hipMemcpyAsync(m_Fc, dFc, sizeof(uint32_t) * eoff[G], hipMemcpyDeviceToDevice, s1));
        BHIP(hipMemcpyAsync(m_Bc, dBc, sizeof(uint32_t) * eoff[G], hipMemcpyDeviceToDevice, s1));
    }
    return finish_upload();
}

// Launch the graphs listed in `idx` with per-graph cfg words and workgroup
// sizes; graphs of each workgroup size go to one persistent launch, the
// launches run concurrently on two streams.  Fills per-graph completion
// times into `lat_ms` (from the earliest graph start).
scc_status BatchEngine::launch(const std::vector<uint32_t> &idx, const std::vector<uint32_t> &cfg,
                               const std::vector<int> &block, double *ms, std::vector<double> *lat_ms,
                               uint32_t *launches)
{
    BHIP(hipSetDevice(dev));
    // group by block size, each group ordered by size (largest first)
    std::vector<uint32_t> order;
    std::vector<std::pair<int, std::pair<uint32_t, uint32_t>>> groups;   // block, (first, count)
    for (int bs : {1024, 512, 256}) {
        std::vector<uint32_t> mine;
        for (uint32_t k = 0; k < idx.size(); k++)
            if (block[k] == bs) mine.push_back(idx[k]);
        if (mine.empty()) continue;
        std::stable_sort(mine.begin(), mine.end(), [&](uint32_t a, uint32_t c) {
            return (uint64_t)V[a] + E[a] > (uint64_t)V[c] + E[c];
        });
        groups.push_back({bs, {(uint32_t)order.size(), (uint32_t)mine.size()}});
        order.insert(order.end(), mine.begin(), mine.end());
    }
    std::vector<uint32_t> cfg_all(G, 0);
    for (uint32_t k = 0; k < idx.size(); k++) cfg_all[idx[k]] = cfg[k];
    BHIP(hipMemcpyAsync(m_cfg, cfg_all.data(), sizeof(uint32_t) * G, hipMemcpyHostToDevice, s1));
    BHIP(hipMemcpyAsync(m_order, order.data(), sizeof(uint32_t) * order.size(), hipMemcpyHostToDevice, s1));
    BHIP(hipMemsetAsync(d.next, 0, 64 * sizeof(uint32_t), s1));
    hipEvent_t e0, ea, eb;
    BHIP(hipEventCreate(&e0));
    BHIP(hipEventCreate(&ea));
    BHIP(hipEventCreate(&eb));
    BHIP(hipEventRecord(e0, s1));
    BHIP(hipStreamWaitEvent(stream2, e0, 0));
    uint32_t nl = 0;
    for (size_t gidx = 0; gidx < groups.size(); gidx++) {
        const int bs = groups[gidx].first;
        const uint32_t first = groups[gidx].second.first, count = groups[gidx].second.second;
        BatchDev bd = d;
        bd.next = d.next + 16 * gidx;      // own 64-byte-apart counter per group
        const uint64_t slots = (uint64_t)num_cus * batch_occupancy(bs);
        const uint32_t grid = (uint32_t)std::min<uint64_t>(count, slots);
        k_batch(gidx % 2 ? stream2 : s1, bd, first, count, bs, grid);
        nl++;
    }
    BHIP(hipEventRecord(ea, s1));
    BHIP(hipEventRecord(eb, stream2));
    BHIP(hipEventSynchronize(ea));
    BHIP(hipEventSynchronize(eb));
    float ma = 0, mb = 0;
    BHIP(hipEventElapsedTime(&ma, e0, ea));
    BHIP(hipEventElapsedTime(&mb, e0, eb));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(ea);
    (void)hipEventDestroy(eb);
    if (ms) *ms = std::max(ma, mb);
    if (launches) *launches = nl;
    if (lat_ms) {
        std::vector<uint64_t> ts(G), te(G);
        BHIP(hipMemcpy(ts.data(), d.tstart, sizeof(uint64_t) * G, hipMemcpyDeviceToHost));
        BHIP(hipMemcpy(te.data(), d.tend, sizeof(uint64_t) * G, hipMemcpyDeviceToHost));
        uint64_t t0 = UINT64_MAX;
        for (uint32_t i : idx) t0 = std::min(t0, ts[i]);
        lat_ms->clear();
        for (uint32_t i : idx) lat_ms->push_back((double)(te[i] - t0) / ticks_per_ms);
    }
    return SCC_OK;
}

static void percentiles(std::vector<double> v, double *p50, double *p99)
{
    if (v.empty()) {
        *p50 = *p99 = 0;
        return;
    }
    std::sort(v.begin(), v.end());
    auto at = [&](double q) {
        size_t k = (size_t)std::ceil(q * v.size());
        return v[std::min(v.size() - 1, k ? k - 1 : 0)];
    };
    *p50 = at(0.50);
    *p99 = at(0.99);
}

scc_status BatchEngine::run(const scc_batch_opts &o, scc_batch_stats *out)
{
    if (!G) return fail(SCC_ERR_NOTREADY, "no batch uploaded");
    if (o.autotuned && !tuned) return fail(SCC_ERR_NOTREADY, "run scc_batch_autotune first");
    if (!o.autotuned) {
        if (o.block != 256 && o.block != 512 && o.block != 1024)
            return fail(SCC_ERR_INVALID, "batch workgroup size %d not in {256, 512, 1024}", o.block);
        if (o.lanes < 1 || o.lanes > 64 || (o.lanes & (o.lanes - 1)))
            return fail(SCC_ERR_INVALID, "lanes %d not a power of two in [1, 64]", o.lanes);
        if (o.algo != 'g' && o.algo != 'h') return fail(SCC_ERR_INVALID, "batch algorithm must be 'g' or 'h'");
    }
    std::vector<uint32_t> idx(G), cfg(G);
    std::vector<int> blk(G);
    for (uint32_t i = 0; i < G; i++) {
        idx[i] = i;
        if (o.autotuned) {
            cfg[i] = best_cfg[bucket[i]];
            blk[i] = best_block[bucket[i]];
        } else {
            cfg[i] = batch_cfg(o.lanes, o.trim1, o.trim2, o.algo == 'h');
            blk[i] = o.block;
        }
    }
    double ms = 0;
    std::vector<double> lat;
    uint32_t nl = 0;
    scc_status s = launch(idx, cfg, blk, &ms, &lat, &nl);
    if (s != SCC_OK) return s;
    std::vector<uint32_t> nscc(G);
    BHIP(hipMemcpy(nscc.data(), d.nscc, sizeof(uint32_t) * G, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < G; i++)
        if (nscc[i] == BATCH_ANOMALY)
            return fail(SCC_ERR_DEVICE, "graph %u: a class finished a round without a pivot", i);
    if (out) {
        std::memset(out, 0, sizeof *out);
        out->graphs = G;
        out->ms = ms;
        out->graphs_per_s = ms > 0 ? G / (ms * 1e-3) : 0;
        percentiles(lat, &out->p50_ms, &out->p99_ms);
        out->vertices = nverts;
        out->arcs = narcs;
        out->launches = nl;
    }
    ran = true;
    return SCC_OK;
}

scc_status BatchEngine::run_serial(Engine &eng, const scc_opts &o, scc_batch_stats *out)
{
    if (!G) return fail(SCC_ERR_NOTREADY, "no batch uploaded");
    BHIP(hipSetDevice(dev));
    std::vector<uint32_t> nscc(G);
    std::vector<double> lat(G);
    uint32_t launches_before = 0;
    (void)launches_before;
    auto t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 0; i < G; i++) {
        scc_status s = eng.upload_device(V[i], E[i], d.Fr + voff[i], d.Fc + eoff[i], d.Br + voff[i],
                                         d.Bc + eoff[i], s1);
        if (s != SCC_OK) return s;
        scc_stats st;
        if ((s = eng.run(o, &st)) != SCC_OK) return s;
        nscc[i] = st.num_sccs;
        BHIP(hipMemcpyAsync(d.lab + voff[i], eng.st.lab, sizeof(uint32_t) * ((size_t)V[i] + 1),
                            hipMemcpyDeviceToDevice, eng.stream));
        BHIP(hipStreamSynchronize(eng.stream));
        lat[i] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    const double ms = lat.empty() ? 0 : lat.back();
    BHIP(hipMemcpy(d.nscc, nscc.data(), sizeof(uint32_t) * G, hipMemcpyHostToDevice));
    if (out) {
        std::memset(out, 0, sizeof *out);
        out->graphs = G;
        out->ms = ms;
        out->graphs_per_s = ms > 0 ? G / (ms * 1e-3) : 0;
        percentiles(lat, &out->p50_ms, &out->p99_ms);
        out->vertices = nverts;
        out->arcs = narcs;
        out->launches = 0;
    }
    ran = true;
    return SCC_OK;
}

// Candidates per bucket: lanes x workgroup size x strategy.  Each candidate
// runs up to `sample` graphs of the bucket in one persistent launch, so the
// graphs share the chip as they do in a full batch; cost = elapsed time of
// that launch (best of two).  A bucket with many graphs is then costed by
// throughput, one with a few large graphs by their latency.
scc_status BatchEngine::autotune(int sample, std::string *report)
{
    if (!G) return fail(SCC_ERR_NOTREADY, "no batch uploaded");
    if (sample < 1) sample = 1;
    struct Cand {
        int lanes, block;
        bool t1, t2, slota;
        const char *name;
    };
    std::vector<Cand> cands;
    for (int lanes : {4, 16, 64})
        for (int blk : {256, 512, 1024}) {
            cands.push_back({lanes, blk, true, true, false, "hong+trim1+trim2"});
            cands.push_back({lanes, blk, true, false, false, "hong+trim1"});
            cands.push_back({lanes, blk, true, false, true, "slota+trim1"});
        }
    best_cfg.assign(NBUCKETS, batch_cfg(16, true, false, true));
    best_block.assign(NBUCKETS, 1024);
    char line[256];
    for (int bk = 0; bk < NBUCKETS; bk++) {
        std::vector<uint32_t> members;
        for (uint32_t i = 0; i < G && (int)members.size() < sample; i++)
            if (bucket[i] == bk) members.push_back(i);
        if (members.empty()) continue;
        double best = 1e300;
        int bi = 0;
        for (size_t c = 0; c < cands.size(); c++) {
            const Cand &cd = cands[c];
            std::vector<uint32_t> cfg(members.size(), batch_cfg(cd.lanes, cd.t1, cd.t2, cd.slota));
            std::vector<int> blk(members.size(), cd.block);
            double cost = 1e300;
            for (int rep = 0; rep < 2; rep++) {
                double ms = 0;
                scc_status s = launch(members, cfg, blk, &ms, nullptr, nullptr);
                if (s != SCC_OK) return s;
                cost = std::min(cost, ms);
            }
            if (cost < best) {
                best = cost;
                bi = (int)c;
            }
        }
        best_cfg[bk] = batch_cfg(cands[bi].lanes, cands[bi].t1, cands[bi].t2, cands[bi].slota);
        best_block[bk] = cands[bi].block;
        if (report) {
            static const char *sz[] = {"2^8-2^10", "2^10-2^12", "2^12-2^14", ">=2^14"};
            static const char *dg[] = {"deg<6", "deg 6-11", "deg>=11"};
            snprintf(line, sizeof line, "bucket V %s %s (%zu graphs): lanes %d, block %d, %s, %.3f ms\n",
                     sz[bk / 3], dg[bk % 3], members.size(), cands[bi].lanes, cands[bi].block,
                     cands[bi].name, best);
            *report += line;
        }
    }
    tuned = true;
    return SCC_OK;
}

scc_status BatchEngine::labels(uint32_t *lab, uint32_t *nscc)
{
    if (!ran) return fail(SCC_ERR_NOTREADY, "no completed batch run");
    BHIP(hipSetDevice(dev));
    if (lab) {
        std::vector<uint32_t> all(voff[G]);
        BHIP(hipMemcpy(all.data(), d.lab, sizeof(uint32_t) * voff[G], hipMemcpyDeviceToHost));
        parallel_chunks(G, [&](uint64_t lo, uint64_t hi, int) {
            for (uint64_t i = lo; i < hi; i++)
                std::memcpy(lab + row_off[i], all.data() + voff[i], sizeof(uint32_t) * ((size_t)V[i] + 1));
        });
    }
    if (nscc) BHIP(hipMemcpy(nscc, d.nscc, sizeof(uint32_t) * G, hipMemcpyDeviceToHost));
    return SCC_OK;
}

} // namespace sccamd

using namespace sccamd;

extern "C" {

scc_status scc_batch_info(scc_ctx *ctx, uint32_t *num_graphs, uint32_t *V, uint32_t *E, uint64_t *row_off,
                          uint64_t *col_off)
{
    clear_error();
    if (!ctx || !num_graphs) return fail(SCC_ERR_INVALID, "null argument");
    if (!ctx->batch) return fail(SCC_ERR_NOTREADY, "no batch uploaded");
    *num_graphs = ctx->batch->info(V, E, row_off, col_off);
    return SCC_OK;
}

scc_status scc_batch_upload(scc_ctx *ctx, const scc_batch_host *b)
{
    clear_error();
    if (!ctx) return fail(SCC_ERR_INVALID, "null context");
    if (!ctx->batch) ctx->batch.reset(new BatchEngine());
    return ctx->batch->upload(ctx->eng.device, ctx->eng.stream, b);
}

scc_status scc_batch_run(scc_ctx *ctx, const scc_batch_opts *opts, scc_batch_stats *stats)
{
    clear_error();
    if (!ctx || !opts) return fail(SCC_ERR_INVALID, "null argument");
    if (!ctx->batch) return fail(SCC_ERR_NOTREADY, "no batch uploaded");
    return ctx->batch->run(*opts, stats);
}

scc_status scc_batch_run_serial(scc_ctx *ctx, const scc_opts *opts, scc_batch_stats *stats)
{
    clear_error();
    if (!ctx) return fail(SCC_ERR_INVALID, "null context");
    if (!ctx->batch) return fail(SCC_ERR_NOTREADY, "no batch uploaded");
    scc_opts o;
    if (opts) o = *opts;
    else scc_opts_default(&o);
    if (!ctx->serial) {
        ctx->serial.reset(new Engine());
        scc_status s = ctx->serial->init(ctx->eng.device);
        if (s != SCC_OK) return s;
    }
    return ctx->batch->run_serial(*ctx->serial, o, stats);
}

scc_status scc_batch_autotune(scc_ctx *ctx, int sample, char *report, size_t report_len)
{
    clear_error();
    if (!ctx) return fail(SCC_ERR_INVALID, "null context");
    if (!ctx->batch) return fail(SCC_ERR_NOTREADY, "no batch uploaded");
    std::string rep;
    scc_status s = ctx->batch->autotune(sample, &rep);
    if (report && report_len) {
        std::strncpy(report, rep.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return s;
}

scc_status scc_batch_labels(scc_ctx *ctx, uint32_t *lab, uint32_t *nscc)
{
    clear_error();
    if (!ctx) return fail(SCC_ERR_INVALID, "null context");
    if (!ctx->batch) return fail(SCC_ERR_NOTREADY, "no batch uploaded");
    return ctx->batch->labels(lab, nscc);
}

} // extern "C"
