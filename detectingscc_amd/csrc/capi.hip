// capi.hip -- extern "C" entry points of libscc_amd.so: context management,
// the device engine, phase entries, and the drop-in drivers with the
// reference's parameters (include/scc.h:49-53) under scc_ names; dropin.cpp
// gives them the reference's own C++ names.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>

#include "batch_engine.h"
#include "engine.h"
#include "host_util.h"
#include "server.h"

using namespace sccamd;

// the server's slot threads stop before the engines they share the device with
scc_ctx::~scc_ctx() { server.reset(); }

#define CTX_CHECK(ctx)                                                                      \
    do {                                                                                    \
        if (!(ctx)) return fail(SCC_ERR_INVALID, "null context");                           \
    } while (0)

extern "C" {

const char *scc_version(void) { return "scc-mi355x 0.1 (gfx950)"; }

void scc_opts_default(scc_opts *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->algo = 'g';
    o->trim1 = 1;
    o->trim2 = 1;
    o->warp = 1;
    o->expand = SCC_EXPAND_AUTO;
}

scc_status scc_ctx_create(int device, scc_ctx **out)
{
    clear_error();
    if (!out) return fail(SCC_ERR_INVALID, "null output pointer");
    *out = nullptr;
    scc_ctx *c = new (std::nothrow) scc_ctx();
    if (!c) return fail(SCC_ERR_NOMEM, "host allocation failed");
    scc_status s = c->eng.init(device);
    if (s != SCC_OK) {
        delete c;
        return s;
    }
    *out = c;
    return SCC_OK;
}

void scc_ctx_destroy(scc_ctx *ctx) { delete ctx; }

scc_status scc_graph_upload(scc_ctx *ctx, uint32_t V, uint32_t E, const uint32_t *Fr,
                            const uint32_t *Fc, const uint32_t *Br, const uint32_t *Bc)
{
    clear_error();
    CTX_CHECK(ctx);
    return ctx->eng.upload(V, E, Fr, Fc, Br, Bc);
}

scc_status scc_partition_upload(scc_ctx *ctx, const uint32_t *part)
{
    clear_error();
    CTX_CHECK(ctx);
    if (!part) return fail(SCC_ERR_INVALID, "null partition");
    return ctx->eng.upload_partition(part);
}

scc_status scc_run(scc_ctx *ctx, const scc_opts *opts, scc_stats *stats)
{
    clear_error();
    CTX_CHECK(ctx);
    scc_opts o;
    if (opts) o = *opts;
    else scc_opts_default(&o);
    return ctx->eng.run(o, stats);
}

scc_status scc_labels(scc_ctx *ctx, uint32_t *lab)
{
    clear_error();
    CTX_CHECK(ctx);
    if (!lab) return fail(SCC_ERR_INVALID, "null label buffer");
    return ctx->eng.labels(lab);
}

scc_status scc_condensation(scc_ctx *ctx, scc_host_graph *dag, uint32_t *node_of)
{
    clear_error();
    CTX_CHECK(ctx);
    if (!dag) return fail(SCC_ERR_INVALID, "null output graph");
    return ctx->eng.condense(dag, node_of);
}

scc_status scc_solve(scc_ctx *ctx, const scc_host_graph *g, const uint32_t *part,
                     const scc_opts *opts, uint32_t *lab, scc_stats *stats)
{
    clear_error();
    CTX_CHECK(ctx);
    if (!g) return fail(SCC_ERR_INVALID, "null graph");
    scc_status s = ctx->eng.upload(g->V, g->E, g->Fr, g->Fc, g->Br, g->Bc);
    if (s == SCC_OK && part) s = ctx->eng.upload_partition(part);
    if (s == SCC_OK) s = scc_run(ctx, opts, stats);
    if (s == SCC_OK && lab) s = ctx->eng.labels(lab);
    return s;
}

scc_status scc_phase_trim1(scc_ctx *ctx, const uint8_t *dead, uint8_t *removed)
{
    clear_error();
    CTX_CHECK(ctx);
    if (!removed) return fail(SCC_ERR_INVALID, "null output");
    return ctx->eng.phase_trim1(dead, removed);
}

scc_status scc_phase_trim2(scc_ctx *ctx, const uint8_t *dead, uint32_t *pair)
{
    clear_error();
    CTX_CHECK(ctx);
    if (!pair) return fail(SCC_ERR_INVALID, "null output");
    return ctx->eng.phase_trim2(dead, pair);
}

scc_status scc_phase_first_pivot(scc_ctx *ctx, const uint8_t *dead, uint32_t *pivot)
{
    clear_error();
    CTX_CHECK(ctx);
    if (!pivot) return fail(SCC_ERR_INVALID, "null output");
    return ctx->eng.phase_first_pivot(dead, pivot);
}

scc_status scc_phase_reach(scc_ctx *ctx, uint32_t src, int dir, int expand, int warp,
                           const uint8_t *dead, uint8_t *out, uint32_t *levels, double *ms)
{
    clear_error();
    CTX_CHECK(ctx);
    if (!out) return fail(SCC_ERR_INVALID, "null output");
    return ctx->eng.phase_reach(src, dir, expand, warp, dead, out, levels, ms);
}

scc_status scc_phase_min_label(scc_ctx *ctx, const uint8_t *dead, uint32_t *out)
{
    clear_error();
    CTX_CHECK(ctx);
    if (!out) return fail(SCC_ERR_INVALID, "null output");
    return ctx->eng.phase_min_label(dead, out);
}

scc_status scc_layout_info(uint32_t V, uint64_t *out, uint32_t n)
{
    clear_error();
    if (!out) return fail(SCC_ERR_INVALID, "null output");
    const uint64_t v[] = {sweep_blocks(V),
                          trim_slot_cap(V),
                          trim_list_words(V),
                          do_blocks(V, 1),
                          do_blocks(V, 2),
                          slot_cap(V, do_blocks(V, 1)),
                          slot_cap(V, do_blocks(V, 2)),
                          queue_words(V),
                          (uint64_t)kSpread,
                          narrow_cap()};
    for (uint32_t i = 0; i < n && i < sizeof v / sizeof v[0]; i++) out[i] = v[i];
    return SCC_OK;
}

scc_status scc_histogram(scc_ctx *ctx, uint64_t *pairs, uint32_t cap, uint32_t *npairs, double *ms)
{
    clear_error();
    CTX_CHECK(ctx);
    if (!npairs || (cap && !pairs)) return fail(SCC_ERR_INVALID, "null output");
    return ctx->eng.histogram(pairs, cap, npairs, ms);
}

scc_status scc_phase_local(scc_ctx *ctx, uint32_t *out)
{
    clear_error();
    CTX_CHECK(ctx);
    if (!out) return fail(SCC_ERR_INVALID, "null output");
    return ctx->eng.phase_local(out);
}

scc_status scc_phase_max_color(scc_ctx *ctx, const uint8_t *dead, uint32_t *out)
{
    clear_error();
    CTX_CHECK(ctx);
    if (!out) return fail(SCC_ERR_INVALID, "null output");
    return ctx->eng.phase_max_color(dead, out);
}

// ------------------------------------------------- drop-in wrappers ---

static std::mutex g_dropin_mu;
static scc_ctx *g_dropin_ctx = nullptr;

// One driver call in the reference's shape: H2D, all phases, D2H of the
// result, then the stdout lines.  "Time" covers the same span as the
// reference's cudaEvent pair (scc_WCC.cu:577-581 -> 951-953): it starts
// before device allocation and ends after the count.
static scc_status dropin(char algo, uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr,
                  uint32_t *Bc, uint32_t *Br, uint32_t *Pr, bool t1, bool t2, int warp)
{
    std::lock_guard<std::mutex> lk(g_dropin_mu);
    auto t0 = std::chrono::steady_clock::now();
    scc_status s = SCC_OK;
    if (!g_dropin_ctx) s = scc_ctx_create(0, &g_dropin_ctx);
    scc_stats st;
    std::memset(&st, 0, sizeof st);
    if (s == SCC_OK) s = g_dropin_ctx->eng.upload(RSize, CSize, Fr, Fc, Br, Bc);
    if (s == SCC_OK && algo == 'd') {
        if (!Pr) s = fail(SCC_ERR_INVALID, "detectSCC needs a partition vector");
        else s = g_dropin_ctx->eng.upload_partition(Pr);
    }
    if (s == SCC_OK) {
        scc_opts o;
        scc_opts_default(&o);
        o.algo = algo;
        o.trim1 = t1;
        o.trim2 = t2;
        o.warp = warp;
        s = g_dropin_ctx->eng.run(o, &st);
    }
    if (s != SCC_OK) return s;   // the message stays in scc_last_error()
    auto t1c = std::chrono::steady_clock::now();
    double ms = std::chrono::duration<double, std::milli>(t1c - t0).count();
    if (algo == 'g') printf("numberOf1Sccs: %u\n", st.num_trim1);
    if (algo == 'd') {
        printf("LOCAL SCC's IDENTIFIED! NODES WITH SAME RANGE VALUES BELONG TO THE SAME SCC!!\n");
        printf("STARTING MERGE!\n");
        printf("MERGING DONE! ^_^\n");
    }
    printf("\nNumber Of Sccs : %u", st.num_sccs);
    printf("\nTime : %f", ms);
    if (algo == 'd') printf("\n");
    fflush(stdout);
    return SCC_OK;
}

scc_status scc_vHong(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc, uint32_t *Br,
          bool t1, bool t2)
{
    return dropin('g', CSize, RSize, Fc, Fr, Bc, Br, nullptr, t1, t2, 1);
}

scc_status scc_wHong(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc, uint32_t *Br,
          bool t1, bool t2, int warpSize)
{
    return dropin('x', CSize, RSize, Fc, Fr, Bc, Br, nullptr, t1, t2, warpSize);
}

scc_status scc_vSlota(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc, uint32_t *Br,
           bool t1, bool t2)
{
    return dropin('h', CSize, RSize, Fc, Fr, Bc, Br, nullptr, t1, t2, 1);
}

scc_status scc_wSlota(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc, uint32_t *Br,
           bool t1, bool t2, int warpSize)
{
    return dropin('y', CSize, RSize, Fc, Fr, Bc, Br, nullptr, t1, t2, warpSize);
}

scc_status scc_detectSCC(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc,
              uint32_t *Br, uint32_t *Pr, bool t1, bool t2)
{
    return dropin('d', CSize, RSize, Fc, Fr, Bc, Br, Pr, t1, t2, 1);
}

} // extern "C"
