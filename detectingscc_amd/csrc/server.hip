// server.hip -- continuous-batching SCC request server.
//
// Requests wait in one queue in arrival order.  Each slot thread owns a
// stream, a BatchEngine (grow-only device buffers) and, for large requests,
// a device-wide Engine.  A free slot takes either the oldest large request
// (V >= solo_vertices), if fewer than slots - 1 slots (at least one) run
// one, and runs it on its device-wide engine; or the small requests in
// arrival order, skipping the large ones, up to `max_batch` and the HBM
// budget left by what the other slots have in flight, as one persistent
// batched launch.  Small requests thus never wait behind large ones for a
// slot, and large requests are taken in arrival order (a burst of them
// takes every slot but one).  Requests that arrive while a batch runs join
// the next one, so the device never waits for a batch to fill.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "batch_engine.h"
#include "engine.h"
#include "host_util.h"
#include "server.h"

namespace sccamd {

uint64_t Server::footprint(uint32_t V, uint32_t E)
{
    // per 64-aligned vertex slot: Fr, Br, cls, lab, aux, pivlist (24 B),
    // 3 pivot keys (24 B), 4 queue words (16 B), 2 visited bits; per arc
    // Fc + Bc; per graph offsets, V, cfg, order, count and two timestamps
    const uint64_t S = (((uint64_t)V + 2 + 63) / 64) * 64;
    return S * 64 + S / 4 + 8ull * E + 48;
}

uint64_t Server::solo_footprint(uint32_t V, uint32_t E)
{
    // the device-wide engine: rows, state, pivot keys, two queue pairs per
    // direction, hub lists and offsets, three bitmaps and byte flags per
    // vertex (under 128 B), 8 B per arc, counters
    return 128ull * ((uint64_t)V + 2) + 8ull * E + 4096;
}

Server::~Server() { stop(); }

scc_status Server::start(uint64_t budget_bytes, uint32_t mb, uint32_t slots, uint32_t solo_v,
                         const scc_batch_opts &o)
{
    if (running) return fail(SCC_ERR_INVALID, "server already running");
    if (!budget_bytes || !mb || !slots || slots > 8)
        return fail(SCC_ERR_INVALID, "server needs budget > 0, max_batch > 0 and 1..8 slots");
    if (o.autotuned) return fail(SCC_ERR_UNSUPPORTED, "the server runs fixed batch options");
    if (o.block != 256 && o.block != 512 && o.block != 1024)
        return fail(SCC_ERR_INVALID, "batch workgroup size %d not in {256, 512, 1024}", o.block);
    if (o.lanes < 1 || o.lanes > 64 || (o.lanes & (o.lanes - 1)))
        return fail(SCC_ERR_INVALID, "lanes %d not a power of two in [1, 64]", o.lanes);
    if (o.algo != 'g' && o.algo != 'h') return fail(SCC_ERR_INVALID, "batch algorithm must be 'g' or 'h'");
    opts = o;
    scc_opts_default(&solo_opts);
    solo_opts.algo = o.algo;
    solo_opts.trim1 = o.trim1;
    solo_opts.trim2 = o.trim2;
    solo_opts.expand = o.algo == 'g' ? SCC_EXPAND_DO : SCC_EXPAND_LB;
    solo_opts.sizes = 0;
    solo_vertices = solo_v ? solo_v : (1u << 16);
    budget = budget_bytes;
    max_batch = mb;
    inflight = peak = 0;
    submitted = completed = rejected = batches = failed = solo_runs = 0;
    latencies.clear();
    busy_ms = 0;
    stopping = false;
    running = true;
    solo_running = 0;
    solo_limit = slots > 1 ? slots - 1 : 1;
    for (uint32_t s = 0; s < slots; s++) workers.emplace_back(&Server::worker, this, (int)s);
    return SCC_OK;
}

scc_status Server::submit(const scc_host_graph *g, uint64_t *id)
{
    if (!running) return fail(SCC_ERR_NOTREADY, "server not started");
    if (!g || !g->Fr || !g->Br || (g->E && (!g->Fc || !g->Bc)))
        return fail(SCC_ERR_INVALID, "incomplete graph");
    if (g->V >= (1u << 30)) return fail(SCC_ERR_INVALID, "V too large for a batch slot");
    const bool solo = g->V >= solo_vertices;
    const uint64_t bytes = solo ? solo_footprint(g->V, g->E) : footprint(g->V, g->E);
    if (bytes > budget) {
        std::lock_guard<std::mutex> lk(mu);
        rejected++;
        return fail(SCC_ERR_BUDGET, "request needs %llu device bytes, budget is %llu",
                    (unsigned long long)bytes, (unsigned long long)budget);
    }
    std::unique_ptr<Req> r(new Req());
    r->V = g->V;
    r->E = g->E;
    r->bytes = bytes;
    r->solo = solo;
    // a bad request must not fail the batch it would share: the copy checks
    // as it goes (one pass, branch-free within a block, so the caller's
    // submit costs about one read and one write of the graph)
    const size_t nr = (size_t)g->V + 2;
    r->Fr.resize(nr);
    r->Br.resize(nr);
    r->Fc.resize(g->E);
    r->Bc.resize(g->E);
    uint32_t down = 0;
    for (size_t v = 0; v < nr; v++) {
        const uint32_t f = g->Fr[v], b = g->Br[v];
        r->Fr[v] = f;
        r->Br[v] = b;
        if (v) down |= (f < g->Fr[v - 1]) | (b < g->Br[v - 1]);
    }
    if (r->Fr[0] || r->Fr[1] || r->Br[0] || r->Br[1] || r->Fr[g->V + 1] != g->E || r->Br[g->V + 1] != g->E)
        return fail(SCC_ERR_INVALID, "row arrays are not in the reference layout");
    if (down) return fail(SCC_ERR_INVALID, "row offsets decrease");
    uint32_t bad = 0;
    const uint32_t hi = g->V - 1;   // column c is valid iff c - 1 <= V - 1 (unsigned)
    for (uint32_t k = 0; k < g->E; k++) {
        const uint32_t f = g->Fc[k], b = g->Bc[k];
        r->Fc[k] = f;
        r->Bc[k] = b;
        bad |= (uint32_t)(f - 1u > hi) | (uint32_t)(b - 1u > hi);
    }
    if (bad) return fail(SCC_ERR_INVALID, "an arc names a vertex outside 1..%u", g->V);
    std::lock_guard<std::mutex> lk(mu);
    r->id = next_id++;
    r->arrival = clock::now();
    if (!submitted) first_arrival = r->arrival;
    submitted++;
    *id = r->id;
    pending.push_back(r->id);
    reqs.emplace(r->id, std::move(r));
    cv_work.notify_one();
    return SCC_OK;
}

// Whether take_locked would take anything: the oldest large request (if no
// slot runs one) or the oldest small request fits in the budget left.
bool Server::can_admit_locked() const
{
    bool large_seen = false, small_seen = false;
    for (uint64_t id : pending) {
        const Req *r = reqs.at(id).get();
        const bool solo_full = solo_running >= solo_limit;
        if (r->solo ? (large_seen || solo_full) : small_seen) continue;
        if (inflight + r->bytes <= budget) return true;
        (r->solo ? large_seen : small_seen) = true;
        if (small_seen && (large_seen || solo_full)) break;
    }
    return false;
}

// What a free slot takes (see the file comment).  The first small request
// that does not fit ends the batch, so small requests keep their order.
void Server::take_locked(std::vector<Req *> &batch, uint64_t &taken)
{
    if (solo_running < solo_limit) {
        for (auto it = pending.begin(); it != pending.end(); ++it) {
            Req *r = reqs.at(*it).get();
            if (!r->solo) continue;
            if (inflight + r->bytes <= budget) {
                batch.push_back(r);
                taken = r->bytes;
                pending.erase(it);
                solo_running++;
                return;
            }
            break;
        }
    }
    for (auto it = pending.begin(); it != pending.end() && batch.size() < max_batch;) {
        Req *r = reqs.at(*it).get();
        if (r->solo) {
            ++it;
            continue;
        }
        if (inflight + taken + r->bytes > budget) break;
        taken += r->bytes;
        batch.push_back(r);
        it = pending.erase(it);
    }
}

void Server::worker(int slot)
{
    (void)slot;
    hipStream_t stream = nullptr;
    std::unique_ptr<BatchEngine> be(new BatchEngine());
    std::unique_ptr<Engine> solo_eng;
    std::string init_err;
    if (hipSetDevice(dev) != hipSuccess || hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess)
        init_err = "cannot create a stream for a server slot";
    std::vector<Req *> batch;
    scc_batch_host h;
    std::vector<uint32_t> hV, hE, Fr, Br, Fc, Bc, nscc, lab;
    std::vector<uint64_t> ro, co;
    for (;;) {
        uint64_t taken = 0;
        batch.clear();
        {
            std::unique_lock<std::mutex> lk(mu);
            // on stop the queue is drained first
            cv_work.wait(lk, [&] { return (stopping && pending.empty()) || can_admit_locked(); });
            if (pending.empty()) break;
            take_locked(batch, taken);
            if (batch.empty()) continue;
            inflight += taken;
            peak = std::max(peak, inflight);
        }
        const uint32_t G = (uint32_t)batch.size();
        ro.assign(G + 1, 0);
        for (uint32_t i = 0; i < G; i++) ro[i + 1] = ro[i] + batch[i]->V + 2;
        lab.resize(ro[G]);
        nscc.resize(G);
        scc_status st = init_err.empty() ? SCC_OK : fail(SCC_ERR_DEVICE, "%s", init_err.c_str());
        if (st == SCC_OK && batch[0]->solo) {
            Req *r = batch[0];
            if (!solo_eng) {
                solo_eng.reset(new Engine());
                st = solo_eng->init(dev);
            }
            if (st == SCC_OK)
                st = solo_eng->upload(r->V, r->E, r->Fr.data(), r->Fc.data(), r->Br.data(), r->Bc.data(),
                                      true);   // checked when it was submitted
            scc_stats sst;
            if (st == SCC_OK) st = solo_eng->run(solo_opts, &sst);
            if (st == SCC_OK) st = solo_eng->labels(lab.data());
            nscc[0] = sst.num_sccs;
        } else if (st == SCC_OK) {
            // pack the batch in the reference layout, one graph after another
            hV.resize(G);
            hE.resize(G);
            co.assign(G + 1, 0);
            for (uint32_t i = 0; i < G; i++) {
                hV[i] = batch[i]->V;
                hE[i] = batch[i]->E;
                co[i + 1] = co[i] + batch[i]->E;
            }
            Fr.resize(ro[G]);
            Br.resize(ro[G]);
            Fc.resize(co[G]);
            Bc.resize(co[G]);
            for (uint32_t i = 0; i < G; i++) {
                std::copy(batch[i]->Fr.begin(), batch[i]->Fr.end(), Fr.begin() + ro[i]);
                std::copy(batch[i]->Br.begin(), batch[i]->Br.end(), Br.begin() + ro[i]);
                std::copy(batch[i]->Fc.begin(), batch[i]->Fc.end(), Fc.begin() + co[i]);
                std::copy(batch[i]->Bc.begin(), batch[i]->Bc.end(), Bc.begin() + co[i]);
            }
            std::memset(&h, 0, sizeof h);
            h.num_graphs = G;
            h.V = hV.data();
            h.E = hE.data();
            h.row_off = ro.data();
            h.col_off = co.data();
            h.Fr = Fr.data();
            h.Br = Br.data();
            h.Fc = Fc.data();
            h.Bc = Bc.data();
            scc_batch_stats bst;
            st = be->upload(dev, stream, &h);
            if (st == SCC_OK) st = be->run(opts, &bst);
            if (st == SCC_OK) st = be->labels(lab.data(), nscc.data());
        }
        const std::string err = st == SCC_OK ? std::string() : std::string(scc_last_error());
        const auto now = clock::now();
        {
            std::lock_guard<std::mutex> lk(mu);
            for (uint32_t i = 0; i < G; i++) {
                Req *r = batch[i];
                r->done = true;
                r->done_at = now;
                r->status = st;
                std::vector<uint32_t>().swap(r->Fr);
                std::vector<uint32_t>().swap(r->Br);
                std::vector<uint32_t>().swap(r->Fc);
                std::vector<uint32_t>().swap(r->Bc);
                if (st == SCC_OK) {
                    r->lab.assign(lab.begin() + ro[i], lab.begin() + ro[i] + r->V + 1);
                    r->nscc = nscc[i];
                    latencies.push_back(std::chrono::duration<double, std::milli>(now - r->arrival).count());
                    completed++;
                } else {
                    r->error = err;
                    failed++;
                }
            }
            last_done = now;
            batches++;
            if (G == 1 && batch[0]->solo) {
                solo_runs++;
                solo_running--;
            }
            inflight -= taken;
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
    be.reset();
    solo_eng.reset();
    if (stream) (void)hipStreamDestroy(stream);
}

scc_status Server::poll(uint64_t id, uint32_t *lab, uint32_t *nscc, double *latency_ms)
{
    std::lock_guard<std::mutex> lk(mu);
    auto it = reqs.find(id);
    if (it == reqs.end()) return fail(SCC_ERR_INVALID, "unknown request %llu", (unsigned long long)id);
    Req *r = it->second.get();
    if (!r->done) return fail(SCC_ERR_NOTREADY, "request %llu not done", (unsigned long long)id);
    scc_status st = r->status;
    if (st == SCC_OK) {
        if (lab) std::memcpy(lab, r->lab.data(), sizeof(uint32_t) * r->lab.size());
        if (nscc) *nscc = r->nscc;
        if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->done_at - r->arrival).count();
    } else {
        fail(st, "%s", r->error.c_str());
    }
    reqs.erase(it);
    return st;
}

scc_status Server::wait_all()
{
    if (!running) return fail(SCC_ERR_NOTREADY, "server not started");
    std::unique_lock<std::mutex> lk(mu);
    cv_done.wait(lk, [&] { return completed + failed == submitted; });
    return SCC_OK;
}

scc_status Server::stats(scc_server_stats *s)
{
    if (!s) return fail(SCC_ERR_INVALID, "null stats");
    std::lock_guard<std::mutex> lk(mu);
    std::memset(s, 0, sizeof *s);
    s->submitted = submitted;
    s->completed = completed;
    s->rejected = rejected;
    s->failed = failed;
    s->batches = batches;
    s->solo = solo_runs;
    s->peak_bytes = peak;
    if (!latencies.empty()) {
        std::vector<double> v = latencies;
        std::sort(v.begin(), v.end());
        auto at = [&](double q) {
            size_t k = (size_t)std::ceil(q * v.size());
            return v[std::min(v.size() - 1, k ? k - 1 : 0)];
        };
        s->p50_ms = at(0.50);
        s->p99_ms = at(0.99);
        double sum = 0;
        for (double x : v) sum += x;
        s->mean_ms = sum / v.size();
        const double span = std::chrono::duration<double>(last_done - first_arrival).count();
        s->graphs_per_s = span > 0 ? completed / span : 0;
    }
    return SCC_OK;
}

scc_status Server::stop()
{
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!running) return SCC_OK;
        stopping = true;
    }
    cv_work.notify_all();
    for (auto &t : workers) t.join();
    workers.clear();
    std::lock_guard<std::mutex> lk(mu);
    running = false;
    stopping = false;
    return SCC_OK;
}

} // namespace sccamd

using namespace sccamd;

extern "C" {

scc_status scc_server_start(scc_ctx *ctx, uint64_t hbm_budget, uint32_t max_batch, uint32_t slots,
                            uint32_t solo_vertices, const scc_batch_opts *opts)
{
    clear_error();
    if (!ctx || !opts) return fail(SCC_ERR_INVALID, "null argument");
    if (!ctx->server) ctx->server.reset(new Server(ctx->eng.device));
    return ctx->server->start(hbm_budget, max_batch, slots, solo_vertices, *opts);
}

scc_status scc_server_submit(scc_ctx *ctx, const scc_host_graph *g, uint64_t *request_id)
{
    clear_error();
    if (!ctx || !request_id) return fail(SCC_ERR_INVALID, "null argument");
    if (!ctx->server) return fail(SCC_ERR_NOTREADY, "server not started");
    return ctx->server->submit(g, request_id);
}

scc_status scc_server_poll(scc_ctx *ctx, uint64_t request_id, uint32_t *lab, uint32_t *nscc,
                           double *latency_ms)
{
    clear_error();
    if (!ctx) return fail(SCC_ERR_INVALID, "null context");
    if (!ctx->server) return fail(SCC_ERR_NOTREADY, "server not started");
    return ctx->server->poll(request_id, lab, nscc, latency_ms);
}

scc_status scc_server_wait_all(scc_ctx *ctx)
{
    clear_error();
    if (!ctx) return fail(SCC_ERR_INVALID, "null context");
    if (!ctx->server) return fail(SCC_ERR_NOTREADY, "server not started");
    return ctx->server->wait_all();
}

scc_status scc_server_stats_get(scc_ctx *ctx, scc_server_stats *stats)
{
    clear_error();
    if (!ctx) return fail(SCC_ERR_INVALID, "null context");
    if (!ctx->server) return fail(SCC_ERR_NOTREADY, "server not started");
    return ctx->server->stats(stats);
}

scc_status scc_server_stop(scc_ctx *ctx)
{
    clear_error();
    if (!ctx) return fail(SCC_ERR_INVALID, "null context");
    if (!ctx->server) return SCC_OK;
    return ctx->server->stop();
}

} // extern "C"
