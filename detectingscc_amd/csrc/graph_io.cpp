// graph_io.cpp -- host data layer: error state, thread helper, GTgraph and
// partition readers, writers, and the multithreaded CSR builder.
//
// Layout produced (reference A1, src/load.cpp:47-113): vertices 1..V, dummy
// vertex 0; Fr/Br have V+2 entries with Fr[0] = Fr[1] = 0 and
// Fr[V+1] = E; Fc is sorted by (src, dst); Bc holds each vertex's
// in-neighbours in ascending order.  Duplicate arcs and self-loops are kept.
#include <cctype>
#include <cstring>
#include <map>

#include "generators.h"
#include "host_util.h"

namespace sccamd {

static thread_local std::string tl_error;

scc_status fail(scc_status st, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    tl_error = buf;
    return st;
}

void clear_error() { tl_error.clear(); }

int host_threads()
{
    const char *e = std::getenv("SCC_HOST_THREADS");
    if (e && std::atoi(e) > 0) return std::atoi(e);
    unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(hc ? hc : 1u, 16u));
}

void parallel_chunks(uint64_t n, const std::function<void(uint64_t, uint64_t, int)> &f,
                     int nthreads)
{
    int nt = nthreads > 0 ? nthreads : host_threads();
    if (n < 4096 || nt == 1) {
        f(0, n, 0);
        return;
    }
    if ((uint64_t)nt > n) nt = (int)n;
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; t++) {
        uint64_t b = n * t / nt, e = n * (t + 1) / nt;
        th.emplace_back([&f, b, e, t] { f(b, e, t); });
    }
    for (auto &x : th) x.join();
}

// Dynamic scheduling over [0, n) in blocks of `grain`.
static void parallel_dynamic(uint64_t n, uint64_t grain,
                             const std::function<void(uint64_t, uint64_t)> &f)
{
    int nt = host_threads();
    if (n <= grain || nt == 1) {
        f(0, n);
        return;
    }
    std::atomic<uint64_t> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
        th.emplace_back([&] {
            for (;;) {
                uint64_t b = next.fetch_add(grain);
                if (b >= n) break;
                f(b, std::min(n, b + grain));
            }
        });
    for (auto &x : th) x.join();
}

// Stable parallel counting sort of n (key, val) pairs by key in [0, K):
// an MSD pass scatters pairs into <= 1024 key buckets (per-thread offsets
// keep input order), then each bucket is counting-sorted on its own by one
// thread.  Outputs the sorted keys/values and rows[k] = #pairs with key < k
// (K+1 entries).  No atomics, so the result is the same for any thread count.
static void stable_sort_pairs(uint64_t n, uint32_t K, const uint32_t *key, const uint32_t *val,
                              uint32_t *key_out, uint32_t *val_out, uint32_t *rows)
{
    uint32_t shift = 0;
    while (((uint64_t)K >> shift) >= 1024) shift++;
    const uint32_t B = (uint32_t)(((uint64_t)K + (1ull << shift) - 1) >> shift);
    const int nt = n < 65536 ? 1 : host_threads();
    std::vector<uint64_t> cnt((size_t)nt * B, 0);
    parallel_chunks(n, [&](uint64_t b, uint64_t e, int t) {
        uint64_t *c = &cnt[(size_t)t * B];
        for (uint64_t i = b; i < e; i++) c[key[i] >> shift]++;
    }, nt);
    std::vector<uint64_t> bstart(B + 1, 0);
    uint64_t off = 0;
    for (uint32_t b = 0; b < B; b++) {
        bstart[b] = off;
        for (int t = 0; t < nt; t++) {
            uint64_t c = cnt[(size_t)t * B + b];
            cnt[(size_t)t * B + b] = off;
            off += c;
        }
    }
    bstart[B] = off;
    uint32_t *tk = xalloc<uint32_t>(n), *tv = xalloc<uint32_t>(n);
    parallel_chunks(n, [&](uint64_t b, uint64_t e, int t) {
        uint64_t *c = &cnt[(size_t)t * B];
        for (uint64_t i = b; i < e; i++) {
            uint64_t p = c[key[i] >> shift]++;
            tk[p] = key[i];
            tv[p] = val[i];
        }
    }, nt);
    rows[0] = 0;
    parallel_dynamic(B, 1, [&](uint64_t b0, uint64_t b1) {
        std::vector<uint64_t> local;
        for (uint64_t b = b0; b < b1; b++) {
            const uint64_t lo = (uint64_t)b << shift;
            const uint64_t width = std::min<uint64_t>(1ull << shift, (uint64_t)K - lo);
            local.assign(width + 1, 0);
            for (uint64_t i = bstart[b]; i < bstart[b + 1]; i++) local[tk[i] - lo + 1]++;
            for (uint64_t k = 0; k < width; k++) {
                rows[lo + k + 1] = (uint32_t)local[k + 1];   // counts; prefixed below
                local[k + 1] += local[k];
            }
            for (uint64_t i = bstart[b]; i < bstart[b + 1]; i++) {
                uint64_t p = bstart[b] + local[tk[i] - lo]++;
                key_out[p] = tk[i];
                val_out[p] = tv[i];
            }
        }
    });
    for (uint64_t k = 1; k <= K; k++) rows[k] += rows[k - 1];
    std::free(tk);
    std::free(tv);
}

// Single-threaded CSR + reverse CSR in the reference layout, for callers
// that already parallelise over many small graphs.  Counting sort by dst,
// then a stable counting sort by src gives rows sorted by (src, dst); a
// stable counting sort of that order by dst gives ascending in-lists.
void csr_build_serial(uint32_t V, uint64_t E, const uint32_t *src, const uint32_t *dst, uint32_t *Fr,
                      uint32_t *Fc, uint32_t *Br, uint32_t *Bc)
{
    std::vector<uint32_t> cnt((size_t)V + 2), s1(E), d1(E);
    auto count_rows = [&](const uint32_t *key, uint32_t *rows) {
        std::fill(cnt.begin(), cnt.end(), 0u);
        for (uint64_t i = 0; i < E; i++) cnt[key[i] + 1]++;
        for (uint32_t k = 1; k <= V + 1; k++) cnt[k] += cnt[k - 1];
        if (rows) std::memcpy(rows, cnt.data(), sizeof(uint32_t) * ((size_t)V + 2));
    };
    count_rows(dst, nullptr);
    for (uint64_t i = 0; i < E; i++) {
        const uint32_t p = cnt[dst[i]]++;
        s1[p] = src[i];
        d1[p] = dst[i];
    }
    count_rows(s1.data(), Fr);
    std::vector<uint32_t> s2(E);
    for (uint64_t i = 0; i < E; i++) {
        const uint32_t p = cnt[s1[i]]++;
        s2[p] = s1[i];
        Fc[p] = d1[i];
    }
    count_rows(Fc, Br);
    for (uint64_t i = 0; i < E; i++) Bc[cnt[Fc[i]]++] = s2[i];
}

} // namespace sccamd

using namespace sccamd;

extern "C" {

const char *scc_last_error(void) { return tl_error.c_str(); }

const char *scc_status_string(int s)
{
    switch (s) {
    case SCC_OK: return "ok";
    case SCC_ERR_INVALID: return "invalid argument";
    case SCC_ERR_IO: return "i/o error";
    case SCC_ERR_PARSE: return "parse error";
    case SCC_ERR_NOMEM: return "out of memory";
    case SCC_ERR_DEVICE: return "device error";
    case SCC_ERR_BUDGET: return "over HBM budget";
    case SCC_ERR_UNSUPPORTED: return "unsupported";
    case SCC_ERR_NOTREADY: return "not ready";
    default: return "unknown status";
    }
}

void scc_host_graph_free(scc_host_graph *g)
{
    if (!g) return;
    std::free(g->Fr);
    std::free(g->Fc);
    std::free(g->Br);
    std::free(g->Bc);
    std::memset(g, 0, sizeof *g);
}

scc_status scc_csr_from_arcs(uint32_t V, uint64_t E, const uint32_t *src, const uint32_t *dst,
                             scc_host_graph *out)
{
    if (!out || (E && (!src || !dst))) return fail(SCC_ERR_INVALID, "null argument");
    if (V >= (1u << 30)) return fail(SCC_ERR_INVALID, "V=%u exceeds 2^30-1", V);
    if (E > 0xFFFFFFFFull) return fail(SCC_ERR_INVALID, "E=%llu exceeds 2^32-1",
                                       (unsigned long long)E);
    std::atomic<int> bad{0};
    parallel_chunks(E, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t i = b; i < e; i++)
            if (src[i] < 1 || src[i] > V || dst[i] < 1 || dst[i] > V) bad = 1;
    });
    if (bad) return fail(SCC_ERR_INVALID, "arc endpoint outside [1, %u]", V);
    scc_host_graph g;
    g.V = V;
    g.E = (uint32_t)E;
    g.Fr = xalloc<uint32_t>((size_t)V + 2);
    g.Br = xalloc<uint32_t>((size_t)V + 2);
    g.Fc = xalloc<uint32_t>(E);
    g.Bc = xalloc<uint32_t>(E);
    if (!g.Fr || !g.Br || !g.Fc || !g.Bc) {
        scc_host_graph_free(&g);
        return fail(SCC_ERR_NOMEM, "host allocation of CSR failed");
    }
    // (src,dst) order = stable sort by dst, then stable sort by src
    // (load.cpp:47); in-lists ascending by source = stable sort of that
    // order by dst (load.cpp:103-112).  Keys are vertex ids in [0, V+1).
    uint32_t *k1 = xalloc<uint32_t>(E), *v1 = xalloc<uint32_t>(E), *k2 = xalloc<uint32_t>(E);
    std::vector<uint32_t> scratch((size_t)V + 2);
    stable_sort_pairs(E, V + 1, dst, src, k1, v1, scratch.data());        // k1 = dst, v1 = src
    stable_sort_pairs(E, V + 1, v1, k1, k2, g.Fc, g.Fr);                   // k2 = src, Fc = dst
    stable_sort_pairs(E, V + 1, g.Fc, k2, k1, g.Bc, g.Br);                 // Bc = src by dst
    // rows[k] = #arcs with key < k (V+2 entries), so row v spans [rows[v], rows[v+1])
    std::free(k1);
    std::free(v1);
    std::free(k2);
    *out = g;
    return SCC_OK;
}

scc_status scc_load_gtgraph(const char *path, scc_host_graph *out)
{
    clear_error();
    if (!path || !out) return fail(SCC_ERR_INVALID, "null argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(SCC_ERR_IO, "cannot open %s", path);
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0) {
        std::fclose(f);
        return fail(SCC_ERR_IO, "cannot size %s", path);
    }
    std::vector<char> buf((size_t)n + 1);
    size_t got = std::fread(buf.data(), 1, (size_t)n, f);
    std::fclose(f);
    if (got != (size_t)n) return fail(SCC_ERR_IO, "short read on %s", path);
    buf[n] = 0;
    const char *p = buf.data(), *end = p + n;
    // load.cpp:32-33: the first 7 lines are a header and are skipped.
    for (int line = 0; line < 7; line++) {
        while (p < end && *p != '\n') p++;
        if (p < end) p++;
    }
    auto ws = [&] {
        while (p < end && std::isspace((unsigned char)*p)) p++;
        return p < end;
    };
    auto num = [&](uint64_t &x) {
        if (!ws() || !std::isdigit((unsigned char)*p)) return false;
        x = 0;
        while (p < end && std::isdigit((unsigned char)*p)) {
            x = x * 10 + (uint64_t)(*p++ - '0');
            if (x > 0xFFFFFFFFull) return false;
        }
        return true;
    };
    // load.cpp:36: "p sp V E" read as char, word, V, E.
    uint64_t V = 0, E = 0;
    if (!ws()) return fail(SCC_ERR_PARSE, "%s: missing problem line", path);
    p++;
    if (!ws()) return fail(SCC_ERR_PARSE, "%s: missing problem line", path);
    while (p < end && !std::isspace((unsigned char)*p)) p++;
    if (!num(V) || !num(E)) return fail(SCC_ERR_PARSE, "%s: bad problem line", path);
    if (V >= (1u << 30)) return fail(SCC_ERR_PARSE, "%s: V=%llu too large", path,
                                     (unsigned long long)V);
    std::vector<uint32_t> s(E), d(E);
    // load.cpp:42-45: E records "a u v w"; the weight is discarded.
    for (uint64_t k = 0; k < E; k++) {
        uint64_t u, v, w;
        if (!ws()) return fail(SCC_ERR_PARSE, "%s: %llu arcs expected, %llu found", path,
                               (unsigned long long)E, (unsigned long long)k);
        p++;
        if (!num(u) || !num(v) || !num(w))
            return fail(SCC_ERR_PARSE, "%s: malformed arc %llu", path, (unsigned long long)k);
        if (u < 1 || u > V || v < 1 || v > V)
            return fail(SCC_ERR_PARSE, "%s: arc %llu (%llu,%llu) outside [1,%llu]", path,
                        (unsigned long long)k, (unsigned long long)u, (unsigned long long)v,
                        (unsigned long long)V);
        s[k] = (uint32_t)u;
        d[k] = (uint32_t)v;
    }
    return scc_csr_from_arcs((uint32_t)V, E, s.data(), d.data(), out);
}

scc_status scc_load_partition(const char *path, uint32_t V, uint32_t *part)
{
    if (!path || !part) return fail(SCC_ERR_INVALID, "null argument");
    FILE *f = std::fopen(path, "r");
    if (!f) return fail(SCC_ERR_IO, "cannot open %s", path);
    part[0] = 0;
    part[V + 1] = 0;
    for (uint32_t k = 0; k < V; k++) {
        unsigned long x;
        if (std::fscanf(f, "%lu", &x) != 1) {
            std::fclose(f);
            return fail(SCC_ERR_PARSE, "%s: %u partition ids expected, %u found", path, V, k);
        }
        if (x > 0xFFFFFFFEul) {
            std::fclose(f);
            return fail(SCC_ERR_PARSE, "%s: partition id %lu too large", path, x);
        }
        part[k + 1] = (uint32_t)x;
    }
    std::fclose(f);
    return SCC_OK;
}

scc_status scc_write_gtgraph(const char *path, const scc_host_graph *g)
{
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(SCC_ERR_IO, "cannot create %s", path);
    std::fprintf(f, "c FILE\t\t\t: %s\n", path);
    std::fprintf(f, "c No. of vertices\t: %u\n", g->V);
    std::fprintf(f, "c No. of directed edges\t: %u\n", g->E);
    std::fprintf(f, "c Max. weight\t\t: 1\n");
    std::fprintf(f, "c Min. weight\t\t: 1\n");
    std::fprintf(f, "c A directed arc from u to v of weight w\n");
    std::fprintf(f, "c is represented below as ' a  u  v  w '\n");
    std::fprintf(f, "p sp %u %u\n", g->V, g->E);
    for (uint32_t u = 1; u <= g->V; u++)
        for (uint32_t i = g->Fr[u]; i < g->Fr[u + 1]; i++) std::fprintf(f, "a %u %u 1\n", u, g->Fc[i]);
    return std::fclose(f) == 0 ? SCC_OK : fail(SCC_ERR_IO, "write to %s failed", path);
}

scc_status scc_write_metis(const char *path, const scc_host_graph *g)
{
    // MetisInput.cpp:115-131: header "V E", then per vertex its out-list
    // followed by its in-list, each id followed by a space.
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(SCC_ERR_IO, "cannot create %s", path);
    std::fprintf(f, "%u %u\n", g->V, g->E);
    for (uint32_t v = 1; v <= g->V; v++) {
        for (uint32_t i = g->Fr[v]; i < g->Fr[v + 1]; i++) std::fprintf(f, "%u ", g->Fc[i]);
        for (uint32_t i = g->Br[v]; i < g->Br[v + 1]; i++) std::fprintf(f, "%u ", g->Bc[i]);
        std::fputc('\n', f);
    }
    return std::fclose(f) == 0 ? SCC_OK : fail(SCC_ERR_IO, "write to %s failed", path);
}

// Binary CSR cache: 8-byte magic, V, E (u32 each), 8 reserved bytes, then
// Fr[V+2], Fc[E], Br[V+2], Bc[E] as little-endian u32.
static const char kCsrMagic[8] = {'S', 'C', 'C', 'C', 'S', 'R', '0', '1'};

scc_status scc_write_csr(const char *path, const scc_host_graph *g)
{
    clear_error();
    if (!path || !g || !g->Fr || !g->Br || (g->E && (!g->Fc || !g->Bc)))
        return fail(SCC_ERR_INVALID, "null argument");
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(SCC_ERR_IO, "cannot create %s", path);
    uint32_t ve[2] = {g->V, g->E};
    uint64_t reserved = 0;
    const size_t rows = (size_t)g->V + 2;
    bool ok = std::fwrite(kCsrMagic, 1, 8, f) == 8 && std::fwrite(ve, 4, 2, f) == 2 &&
              std::fwrite(&reserved, 8, 1, f) == 1 && std::fwrite(g->Fr, 4, rows, f) == rows &&
              std::fwrite(g->Fc, 4, g->E, f) == g->E && std::fwrite(g->Br, 4, rows, f) == rows &&
              std::fwrite(g->Bc, 4, g->E, f) == g->E;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? SCC_OK : fail(SCC_ERR_IO, "write to %s failed", path);
}

scc_status scc_load_csr(const char *path, scc_host_graph *out)
{
    clear_error();
    if (!path || !out) return fail(SCC_ERR_INVALID, "null argument");
    std::memset(out, 0, sizeof *out);
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(SCC_ERR_IO, "cannot open %s", path);
    char magic[8];
    uint32_t ve[2];
    uint64_t reserved;
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kCsrMagic, 8) != 0 ||
        std::fread(ve, 4, 2, f) != 2 || std::fread(&reserved, 8, 1, f) != 1) {
        std::fclose(f);
        return fail(SCC_ERR_PARSE, "%s: not a binary CSR file", path);
    }
    const uint32_t V = ve[0], E = ve[1];
    if (V >= (1u << 30)) {
        std::fclose(f);
        return fail(SCC_ERR_PARSE, "%s: V = %u too large", path, V);
    }
    const size_t rows = (size_t)V + 2;
    scc_host_graph g;
    g.V = V;
    g.E = E;
    g.Fr = xalloc<uint32_t>(rows);
    g.Br = xalloc<uint32_t>(rows);
    g.Fc = xalloc<uint32_t>(E ? E : 1);
    g.Bc = xalloc<uint32_t>(E ? E : 1);
    if (!g.Fr || !g.Br || !g.Fc || !g.Bc) {
        std::fclose(f);
        scc_host_graph_free(&g);
        return fail(SCC_ERR_NOMEM, "host allocation failed for %s", path);
    }
    const bool ok = std::fread(g.Fr, 4, rows, f) == rows && std::fread(g.Fc, 4, E, f) == E &&
                    std::fread(g.Br, 4, rows, f) == rows && std::fread(g.Bc, 4, E, f) == E;
    std::fclose(f);
    if (!ok) {
        scc_host_graph_free(&g);
        return fail(SCC_ERR_PARSE, "%s: file shorter than its header says", path);
    }
    // the reference layout: rows from 0,0 to E, monotone; rows sorted and ids
    // in [1, V]; the reverse CSR's in-degrees match the forward columns
    std::atomic<int> bad{0};
    if (g.Fr[0] || g.Fr[1] || g.Br[0] || g.Br[1] || g.Fr[V + 1] != E || g.Br[V + 1] != E) bad = 1;
    if (!bad) {
        parallel_chunks((uint64_t)V + 1, [&](uint64_t lo, uint64_t hi, int) {
            for (uint64_t v = lo; v < hi && !bad; v++) {
                if (g.Fr[v] > g.Fr[v + 1] || g.Br[v] > g.Br[v + 1]) { bad = 1; break; }
                for (uint32_t i = g.Fr[v]; i < g.Fr[v + 1]; i++)
                    if (g.Fc[i] < 1 || g.Fc[i] > V || (i > g.Fr[v] && g.Fc[i] < g.Fc[i - 1])) { bad = 1; break; }
                for (uint32_t i = g.Br[v]; i < g.Br[v + 1]; i++)
                    if (g.Bc[i] < 1 || g.Bc[i] > V || (i > g.Br[v] && g.Bc[i] < g.Bc[i - 1])) { bad = 1; break; }
            }
        });
    }
    if (!bad) {
        std::vector<uint32_t> indeg((size_t)V + 2, 0);
        for (uint32_t i = 0; i < E; i++) indeg[g.Fc[i]]++;
        for (uint32_t v = 1; v <= V && !bad; v++)
            if (g.Br[v + 1] - g.Br[v] != indeg[v]) bad = 1;
    }
    if (bad) {
        scc_host_graph_free(&g);
        return fail(SCC_ERR_PARSE, "%s: arrays are not a CSR + reverse CSR in the reference layout", path);
    }
    *out = g;
    return SCC_OK;
}

scc_status scc_load_graph(const char *path, scc_host_graph *out)
{
    clear_error();
    if (!path || !out) return fail(SCC_ERR_INVALID, "null argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(SCC_ERR_IO, "cannot open %s", path);
    char magic[8] = {0};
    const size_t n = std::fread(magic, 1, 8, f);
    std::fclose(f);
    if (n == 8 && std::memcmp(magic, kCsrMagic, 8) == 0) return scc_load_csr(path, out);
    return scc_load_gtgraph(path, out);
}

scc_status scc_size_histogram(uint32_t V, const uint32_t *lab, uint64_t *pairs, uint32_t cap,
                              uint32_t *npairs)
{
    if (!lab || !npairs || (cap && !pairs)) return fail(SCC_ERR_INVALID, "null argument");
    std::vector<uint32_t> size((size_t)V + 1, 0);
    for (uint32_t v = 1; v <= V; v++) {
        if (lab[v] < 1 || lab[v] > v) return fail(SCC_ERR_INVALID, "label %u of vertex %u is not canonical", lab[v], v);
        size[lab[v]]++;
    }
    std::map<uint32_t, uint64_t> h;
    for (uint32_t v = 1; v <= V; v++)
        if (size[v]) h[size[v]]++;
    uint32_t k = 0;
    for (auto &kv : h) {
        if (k < cap) {
            pairs[2 * k] = kv.first;
            pairs[2 * k + 1] = kv.second;
        }
        k++;
    }
    *npairs = k;
    return SCC_OK;
}

} // extern "C"
