// generators.cpp -- seeded synthetic digraphs for tests, benchmarks, the
// batched mix and the serving stream.  Every random choice is a pure
// function rnd(seed, stream, index), so results do not depend on the
// number of host threads.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <unordered_set>

#include "generators.h"
#include "host_util.h"

namespace sccamd {

void er_arcs(uint32_t V, uint64_t draws, uint64_t seed, std::vector<uint32_t> &s,
             std::vector<uint32_t> &d, int threads)
{
    size_t base = s.size();
    std::vector<uint32_t> u(draws), v(draws);
    parallel_chunks(draws, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t k = b; k < e; k++) {
            u[k] = 1 + (uint32_t)below(rnd(seed, 1, 2 * k), V);
            v[k] = 1 + (uint32_t)below(rnd(seed, 1, 2 * k + 1), V);
        }
    }, threads);
    s.reserve(base + draws);
    d.reserve(base + draws);
    for (uint64_t k = 0; k < draws; k++)
        if (u[k] != v[k]) {
            s.push_back(u[k]);
            d.push_back(v[k]);
        }
}

void planted_cycles(uint32_t V, uint32_t cycles, uint32_t cmin, uint32_t cmax, uint64_t seed,
                    std::vector<uint32_t> &s, std::vector<uint32_t> &d)
{
    if (cmax > V) cmax = V;
    if (cmin < 2) cmin = 2;
    if (cmin > cmax) return;
    std::vector<uint32_t> cyc;
    std::unordered_set<uint32_t> used;
    for (uint32_t j = 0; j < cycles; j++) {
        uint32_t L = cmin + (uint32_t)below(rnd(seed, 2, j), (uint64_t)(cmax - cmin + 1));
        cyc.clear();
        used.clear();
        for (uint64_t t = 0; cyc.size() < L; t++) {
            uint32_t x = 1 + (uint32_t)below(rnd(seed, 3 + (uint64_t)j, t), V);
            if (used.insert(x).second) cyc.push_back(x);
        }
        for (uint32_t i = 0; i < L; i++) {
            s.push_back(cyc[i]);
            d.push_back(cyc[(i + 1) % L]);
        }
    }
}

RmatParams rmat_params(uint32_t scale, uint32_t vcap, uint64_t E, double a, double b, double c)
{
    RmatParams P;
    P.ta = (uint32_t)std::lround(a * 65536.0);
    P.tab = (uint32_t)std::lround((a + b) * 65536.0);
    P.tabc = (uint32_t)std::lround((a + b + c) * 65536.0);
    P.scale = scale;
    P.words = (scale + 3) / 4;
    P.vcap = vcap;
    P.E = E;
    return P;
}

// R-MAT over [1, 2^scale] with 16-bit quadrant thresholds per level; arcs
// with an endpoint above vcap are redrawn (counter advanced), which gives
// R-MAT graphs of any vertex count.  With `permute`, both endpoints go
// through the seeded permutation of [0, vcap).
void rmat_arcs(uint32_t scale, uint32_t vcap, uint64_t E, double a, double b, double c,
               uint64_t seed, bool permute, std::vector<uint32_t> &s, std::vector<uint32_t> &d, int threads)
{
    const RmatParams P = rmat_params(scale, vcap, E, a, b, c);
    const SeededPerm perm = make_perm(vcap, seed);
    size_t base = s.size();
    s.resize(base + E);
    d.resize(base + E);
    parallel_chunks(E, [&](uint64_t lo, uint64_t hi, int) {
        for (uint64_t i = lo; i < hi; i++) {
            uint32_t r, q;
            rmat_arc(P, seed, i, r, q);
            if (permute) {
                r = perm_at(perm, r);
                q = perm_at(perm, q);
            }
            s[base + i] = r + 1;
            d[base + i] = q + 1;
        }
    }, threads);
}

// Graph `index` of the batched mix (the many-graph workload): V log-uniform in
// [2^lgmin, 2^lgmax], mean out-degree uniform in [dmin, dmax]; 40% ER,
// 40% R-MAT (a=.57 b=.19 c=.19, ids above V redrawn, seeded permutation),
// 20% planted directed cycles of length 2..512 carrying about half of the
// arcs plus ER arcs for the rest.  Single-threaded: callers run many
// graphs in parallel.
MixParams mix_params(uint64_t seed, uint32_t index, uint32_t lgmin, uint32_t lgmax, double dmin, double dmax)
{
    MixParams m;
    m.gs = mix64(seed * 0x9E3779B97F4A7C15ull + index + 1);
    const double lg = lgmin + unit(rnd(m.gs, 40, 0)) * (double)(lgmax - lgmin);
    m.V = (uint32_t)std::llround(std::pow(2.0, lg));
    if (m.V < 2) m.V = 2;
    const double deg = dmin + unit(rnd(m.gs, 40, 1)) * (dmax - dmin);
    m.E = (uint64_t)std::llround(deg * m.V);
    const double t = unit(rnd(m.gs, 40, 2));
    m.kind = t < 0.4 ? 0 : (t < 0.8 ? 1 : 2);
    return m;
}

// About half of a planted-cycle graph's arcs are on cycles of mean length
// 257 (lengths uniform in 2..512, capped at V).
uint32_t mix_cycles(const MixParams &m) { return (uint32_t)std::max<uint64_t>(1, m.E / (2 * 257)); }

uint32_t mix_cycle_len(const MixParams &m, uint32_t j)
{
    const uint32_t cmax = std::min<uint32_t>(512, m.V);   // planted_cycles' clamp (V >= 2)
    return 2 + (uint32_t)below(rnd(m.gs, 2, j), (uint64_t)(cmax - 2 + 1));
}

MixGraph mix_graph(uint64_t seed, uint32_t index, uint32_t lgmin, uint32_t lgmax, double dmin, double dmax)
{
    const MixParams p = mix_params(seed, index, lgmin, lgmax, dmin, dmax);
    MixGraph m;
    m.V = p.V;
    m.kind = p.kind;
    const uint64_t gs = p.gs, E = p.E;
    if (m.kind == 0) {
        er_arcs(m.V, E, gs, m.s, m.d, 1);
    } else if (m.kind == 1) {
        uint32_t scale = 1;
        while ((1ull << scale) < m.V) scale++;
        rmat_arcs(scale, m.V, E, 0.57, 0.19, 0.19, gs, true, m.s, m.d, 1);
    } else {
        planted_cycles(m.V, mix_cycles(p), 2, 512, gs, m.s, m.d);
        const uint64_t have = m.s.size();
        if (have < E) er_arcs(m.V, E - have, gs ^ kMixFillSalt, m.s, m.d, 1);
    }
    return m;
}

MixGraph serving_graph(uint64_t seed, uint32_t index, double median_v, double sigma, uint32_t vcap,
                       double degree)
{
    MixGraph m;
    const uint64_t gs = mix64(seed * 0xD1B54A32D192ED03ull + index + 1);
    // lognormal V: standard normal by Box-Muller from two uniforms in (0, 1]
    const double u1 = 1.0 - unit(rnd(gs, 41, 0)), u2 = unit(rnd(gs, 41, 1));
    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    const double v = (double)std::llround(median_v * std::exp(sigma * z));
    m.V = (uint32_t)std::min<double>(vcap, std::max<double>(2.0, v));
    const uint64_t E = (uint64_t)std::llround(degree * m.V);
    m.kind = unit(rnd(gs, 41, 2)) < 0.5 ? 0 : 1;
    if (m.kind == 0) {
        er_arcs(m.V, E, gs, m.s, m.d, 1);
    } else {
        uint32_t scale = 1;
        while ((1ull << scale) < m.V) scale++;
        rmat_arcs(scale, m.V, E, 0.57, 0.19, 0.19, gs, true, m.s, m.d, 1);
    }
    return m;
}

void grid_arcs(uint32_t side, double p, uint64_t seed, std::vector<uint32_t> &s,
               std::vector<uint32_t> &d)
{
    const uint64_t V = (uint64_t)side * side;
    std::vector<uint8_t> mask(V);
    parallel_chunks(V, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t x = b; x < e; x++) {
            uint64_t r = x / side, c = x % side, id = x + 1;
            uint8_t m = 0;
            // directions in ascending target order: N, W, E, S
            if (r > 0 && bernoulli(rnd(seed, 20, id * 4 + 0), p)) m |= 1;
            if (c > 0 && bernoulli(rnd(seed, 20, id * 4 + 1), p)) m |= 2;
            if (c + 1 < side && bernoulli(rnd(seed, 20, id * 4 + 2), p)) m |= 4;
            if (r + 1 < side && bernoulli(rnd(seed, 20, id * 4 + 3), p)) m |= 8;
            mask[x] = m;
        }
    });
    uint64_t n = 0;
    for (uint64_t x = 0; x < V; x++) n += __builtin_popcount(mask[x]);
    s.reserve(s.size() + n);
    d.reserve(d.size() + n);
    for (uint64_t x = 0; x < V; x++) {
        uint32_t id = (uint32_t)x + 1;
        if (mask[x] & 1) { s.push_back(id); d.push_back(id - side); }
        if (mask[x] & 2) { s.push_back(id); d.push_back(id - 1); }
        if (mask[x] & 4) { s.push_back(id); d.push_back(id + 1); }
        if (mask[x] & 8) { s.push_back(id); d.push_back(id + side); }
    }
}

} // namespace sccamd

using namespace sccamd;

static scc_status finish(uint32_t V, std::vector<uint32_t> &s, std::vector<uint32_t> &d,
                         scc_host_graph *out)
{
    if (s.size() > 0xFFFFFFFFull) return fail(SCC_ERR_INVALID, "too many arcs");
    return scc_csr_from_arcs(V, s.size(), s.data(), d.data(), out);
}

extern "C" {

scc_status scc_gen_er(uint32_t V, uint64_t draws, uint32_t cycles, uint32_t cmin, uint32_t cmax,
                      uint64_t seed, scc_host_graph *out)
{
    clear_error();
    if (!out || V < 1 || V >= (1u << 30)) return fail(SCC_ERR_INVALID, "bad ER arguments");
    std::vector<uint32_t> s, d;
    er_arcs(V, draws, seed, s, d);
    planted_cycles(V, cycles, cmin, cmax, seed, s, d);
    return finish(V, s, d, out);
}

scc_status scc_gen_rmat(uint32_t scale, uint64_t E, double a, double b, double c, uint64_t seed,
                        scc_host_graph *out)
{
    clear_error();
    if (!out || scale < 1 || scale > 29 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return fail(SCC_ERR_INVALID, "bad R-MAT arguments");
    uint32_t V = 1u << scale;
    std::vector<uint32_t> s, d;
    rmat_arcs(scale, V, E, a, b, c, seed, true, s, d);
    return finish(V, s, d, out);
}

scc_status scc_seeded_permutation(uint32_t n, uint64_t seed, uint32_t *out)
{
    clear_error();
    if (!out) return fail(SCC_ERR_INVALID, "null output");
    const SeededPerm p = make_perm(n, seed);
    parallel_chunks(n, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t i = b; i < e; i++) out[i] = perm_at(p, (uint32_t)i);
    });
    return SCC_OK;
}

scc_status scc_gen_grid(uint32_t side, double p, uint64_t seed, scc_host_graph *out)
{
    clear_error();
    if (!out || side < 1 || (uint64_t)side * side >= (1u << 30) || p < 0 || p > 1)
        return fail(SCC_ERR_INVALID, "bad grid arguments");
    std::vector<uint32_t> s, d;
    grid_arcs(side, p, seed, s, d);
    return finish(side * side, s, d, out);
}

void scc_batch_host_free(scc_batch_host *b)
{
    if (!b) return;
    std::free(b->V);
    std::free(b->E);
    std::free(b->row_off);
    std::free(b->col_off);
    std::free(b->Fr);
    std::free(b->Fc);
    std::free(b->Br);
    std::free(b->Bc);
    std::free(b->kind);
    std::memset(b, 0, sizeof *b);
}

} // extern "C"

// Lays out G generated graphs (make(i) -> MixGraph) as one host batch: graphs
// are generated in parallel, then each is sorted into its CSR slice.
template <class Make> static scc_status assemble_batch(uint32_t G, Make make, scc_batch_host *out)
{
    std::vector<MixGraph> graphs(G);
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> th;
    const int nt = host_threads();
    for (int t = 0; t < nt; t++)
        th.emplace_back([&] {
            for (uint32_t i; (i = next.fetch_add(1)) < G;) graphs[i] = make(i);
        });
    for (auto &x : th) x.join();
    scc_batch_host b;
    std::memset(&b, 0, sizeof b);
    b.num_graphs = G;
    b.V = xalloc<uint32_t>(G);
    b.E = xalloc<uint32_t>(G);
    b.kind = xalloc<uint8_t>(G);
    b.row_off = xalloc<uint64_t>((size_t)G + 1);
    b.col_off = xalloc<uint64_t>((size_t)G + 1);
    if (!b.V || !b.E || !b.kind || !b.row_off || !b.col_off) {
        scc_batch_host_free(&b);
        return fail(SCC_ERR_NOMEM, "host allocation of the batch failed");
    }
    b.row_off[0] = b.col_off[0] = 0;
    for (uint32_t i = 0; i < G; i++) {
        if (graphs[i].s.size() > 0xFFFFFFFFull) {
            scc_batch_host_free(&b);
            return fail(SCC_ERR_INVALID, "graph %u has too many arcs", i);
        }
        b.V[i] = graphs[i].V;
        b.E[i] = (uint32_t)graphs[i].s.size();
        b.kind[i] = graphs[i].kind;
        b.row_off[i + 1] = b.row_off[i] + graphs[i].V + 2;
        b.col_off[i + 1] = b.col_off[i] + b.E[i];
    }
    b.Fr = xalloc<uint32_t>(b.row_off[G]);
    b.Br = xalloc<uint32_t>(b.row_off[G]);
    b.Fc = xalloc<uint32_t>(b.col_off[G]);
    b.Bc = xalloc<uint32_t>(b.col_off[G]);
    if (!b.Fr || !b.Br || !b.Fc || !b.Bc) {
        scc_batch_host_free(&b);
        return fail(SCC_ERR_NOMEM, "host allocation of the batch failed");
    }
    next = 0;
    th.clear();
    for (int t = 0; t < nt; t++)
        th.emplace_back([&] {
            for (uint32_t i; (i = next.fetch_add(1)) < G;) {
                MixGraph &m = graphs[i];
                csr_build_serial(m.V, m.s.size(), m.s.data(), m.d.data(), b.Fr + b.row_off[i],
                                 b.Fc + b.col_off[i], b.Br + b.row_off[i], b.Bc + b.col_off[i]);
                std::vector<uint32_t>().swap(m.s);
                std::vector<uint32_t>().swap(m.d);
            }
        });
    for (auto &x : th) x.join();
    *out = b;
    return SCC_OK;
}

extern "C" {

scc_status scc_gen_batch_mix(uint32_t num_graphs, uint32_t lgmin, uint32_t lgmax, double dmin,
                             double dmax, uint64_t seed, scc_batch_host *out)
{
    clear_error();
    if (!out || !num_graphs || lgmin < 1 || lgmax < lgmin || lgmax > 24 || dmin < 0 || dmax < dmin)
        return fail(SCC_ERR_INVALID, "bad batch-mix arguments");
    return assemble_batch(
        num_graphs, [&](uint32_t i) { return mix_graph(seed, i, lgmin, lgmax, dmin, dmax); }, out);
}

scc_status scc_gen_serving_mix(uint32_t num_graphs, double median_v, double sigma, uint32_t vcap,
                               double degree, uint64_t seed, scc_batch_host *out)
{
    clear_error();
    if (!out || !num_graphs || median_v < 2 || sigma < 0 || vcap < 2 || vcap >= (1u << 30) || degree < 0)
        return fail(SCC_ERR_INVALID, "bad serving-mix arguments");
    return assemble_batch(
        num_graphs, [&](uint32_t i) { return serving_graph(seed, i, median_v, sigma, vcap, degree); }, out);
}

scc_status scc_gen_communities(uint32_t V, uint32_t K, double deg, double inter, uint64_t seed,
                               scc_host_graph *out, uint32_t *part)
{
    clear_error();
    if (!out || !part || V < 1 || K < 1 || K > V || V >= (1u << 30) || deg < 0 || inter < 0 ||
        inter > 1)
        return fail(SCC_ERR_INVALID, "bad community arguments");
    const SeededPerm perm = make_perm(V, seed);   // position -> vertex-1
    // community c holds positions [c*V/K, (c+1)*V/K)
    auto lo = [&](uint64_t c) { return (uint32_t)(c * V / K); };
    part[0] = 0;
    part[V + 1] = 0;
    parallel_chunks(V, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t v = b; v < e; v++) {
            const uint32_t pos = perm_inv(perm, (uint32_t)v);
            uint32_t c = (uint32_t)(((uint64_t)pos * K) / V);
            while (c + 1 < K && pos >= lo(c + 1)) c++;   // integer rounding can misplace the boundary
            while (c > 0 && pos < lo(c)) c--;
            part[v + 1] = c;
        }
    });
    const uint64_t E = (uint64_t)std::llround(deg * V);
    std::vector<uint32_t> s(E), d(E);
    parallel_chunks(E, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t i = b; i < e; i++) {
            uint32_t u = (uint32_t)below(rnd(seed, 30, i), V);
            uint32_t v;
            if (bernoulli(rnd(seed, 31, i), inter)) {
                v = (uint32_t)below(rnd(seed, 32, i), V);
            } else {
                uint32_t c = part[u + 1];
                uint32_t a0 = lo(c), n = lo(c + 1) - a0;
                v = perm_at(perm, a0 + (uint32_t)below(rnd(seed, 32, i), n));
            }
            s[i] = u + 1;
            d[i] = v + 1;
        }
    });
    return finish(V, s, d, out);
}

} // extern "C"
