/*
 * oracle_core.c -- CPU oracle: graph I/O restatement and ground-truth SCC.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Plain C, single-threaded.
 */
#include "oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>

void orc_free(void *p) { free(p); }

/* ---------------------------------------------------------------- I/O --- */

/* Sort key for (src,dst) pairs: load.cpp:8-10 compares first, then second. */
static int cmp_pair(const void *a, const void *b)
{
    const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/*
 * Build the reference CSR layout from an arc list (load.cpp:47-113):
 *   - arcs sorted by (src, dst)                      (load.cpp:47)
 *   - Fr[0]=Fr[1]=0, Fr[v] = #arcs with src < v      (load.cpp:55-77)
 *   - Br[v] = #arcs with dst < v                     (load.cpp:89-101)
 *   - Bc filled by a scan of Fc in order, so each in-list is ascending
 *     by source                                      (load.cpp:103-112)
 */
void orc_build_csr(uint32_t V, uint32_t E, const uint32_t *src, const uint32_t *dst,
                   uint32_t *Fr, uint32_t *Fc, uint32_t *Br, uint32_t *Bc)
{
    uint64_t *arcs = (uint64_t *)malloc(sizeof(uint64_t) * (E ? E : 1));
    for (uint32_t i = 0; i < E; i++)
        arcs[i] = ((uint64_t)src[i] << 32) | dst[i];
    qsort(arcs, E, sizeof(uint64_t), cmp_pair);
    memset(Fr, 0, sizeof(uint32_t) * ((size_t)V + 2));
    memset(Br, 0, sizeof(uint32_t) * ((size_t)V + 2));
    for (uint32_t i = 0; i < E; i++) {
        uint32_t u = (uint32_t)(arcs[i] >> 32), v = (uint32_t)arcs[i];
        Fc[i] = v;
        Fr[u + 1]++;
        Br[v + 1]++;
    }
    for (uint32_t v = 1; v <= V + 1; v++) {
        Fr[v] += Fr[v - 1];
        Br[v] += Br[v - 1];
    }
    uint32_t *fill = (uint32_t *)calloc((size_t)V + 2, sizeof(uint32_t));
    for (uint32_t u = 1; u <= V; u++)
        for (uint32_t i = Fr[u]; i < Fr[u + 1]; i++) {
            uint32_t v = Fc[i];
            Bc[Br[v] + fill[v]++] = u;
        }
    free(fill);
    free(arcs);
}

/* whitespace-token reader over a whole file */
typedef struct { const char *p, *end; } tok_t;

static int tok_skip_ws(tok_t *t)
{
    while (t->p < t->end && isspace((unsigned char)*t->p)) t->p++;
    return t->p < t->end;
}

static int tok_char(tok_t *t, char *c)
{
    if (!tok_skip_ws(t)) return 0;
    *c = *t->p++;
    return 1;
}

static int tok_word(tok_t *t)
{
    if (!tok_skip_ws(t)) return 0;
    while (t->p < t->end && !isspace((unsigned char)*t->p)) t->p++;
    return 1;
}

static int tok_u64(tok_t *t, uint64_t *x)
{
    if (!tok_skip_ws(t)) return 0;
    if (!isdigit((unsigned char)*t->p)) return 0;
    uint64_t v = 0;
    while (t->p < t->end && isdigit((unsigned char)*t->p)) {
        v = v * 10 + (uint64_t)(*t->p - '0');
        if (v > 0xFFFFFFFFull) return 0;
        t->p++;
    }
    *x = v;
    return 1;
}

/*
 * GTgraph reader (load.cpp:12-113): skip exactly 7 lines (load.cpp:32-33),
 * read "p sp V E" (load.cpp:36), then E records "a u v w" with the weight
 * discarded (load.cpp:42-45).  Unlike the reference, ids outside [1,V] and
 * short files are reported as errors (-3) instead of being undefined.
 */
int orc_load_gtgraph(const char *path, uint32_t *oV, uint32_t *oE,
                     uint32_t **oFr, uint32_t **oFc, uint32_t **oBr, uint32_t **oBc)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -2;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *buf = (char *)malloc((size_t)n + 1);
    if (fread(buf, 1, (size_t)n, f) != (size_t)n) { fclose(f); free(buf); return -2; }
    fclose(f);
    buf[n] = 0;
    tok_t t = { buf, buf + n };
    for (int line = 0; line < 7; line++) {
        while (t.p < t.end && *t.p != '\n') t.p++;
        if (t.p < t.end) t.p++;
    }
    char c;
    uint64_t V, E, u, v, w;
    if (!tok_char(&t, &c) || !tok_word(&t) || !tok_u64(&t, &V) || !tok_u64(&t, &E)) {
        free(buf);
        return -3;
    }
    uint32_t *src = (uint32_t *)malloc(sizeof(uint32_t) * (E ? E : 1));
    uint32_t *dst = (uint32_t *)malloc(sizeof(uint32_t) * (E ? E : 1));
    for (uint64_t k = 0; k < E; k++) {
        if (!tok_char(&t, &c) || !tok_u64(&t, &u) || !tok_u64(&t, &v) || !tok_u64(&t, &w)
            || u < 1 || u > V || v < 1 || v > V) {
            free(src); free(dst); free(buf);
            return -3;
        }
        src[k] = (uint32_t)u;
        dst[k] = (uint32_t)v;
    }
    free(buf);
    uint32_t *Fr = (uint32_t *)malloc(sizeof(uint32_t) * (V + 2));
    uint32_t *Br = (uint32_t *)malloc(sizeof(uint32_t) * (V + 2));
    uint32_t *Fc = (uint32_t *)malloc(sizeof(uint32_t) * (E ? E : 1));
    uint32_t *Bc = (uint32_t *)malloc(sizeof(uint32_t) * (E ? E : 1));
    orc_build_csr((uint32_t)V, (uint32_t)E, src, dst, Fr, Fc, Br, Bc);
    free(src);
    free(dst);
    *oV = (uint32_t)V; *oE = (uint32_t)E;
    *oFr = Fr; *oFc = Fc; *oBr = Br; *oBc = Bc;
    return 0;
}

/* Partition vector (load.cpp:117-130): V whitespace-separated ids, Pr[k+1]. */
int orc_load_partition(const char *path, uint32_t V, uint32_t *part)
{
    FILE *f = fopen(path, "r");
    if (!f) return -2;
    part[0] = 0;
    part[V + 1] = 0;
    for (uint32_t k = 0; k < V; k++) {
        unsigned long x;
        if (fscanf(f, "%lu", &x) != 1) { fclose(f); return -3; }
        part[k + 1] = (uint32_t)x;
    }
    fclose(f);
    return 0;
}

/* ------------------------------------------------------ ground truth --- */

/*
 * Iterative Tarjan restricted to vertices with dead[v]==0 (dead may be NULL).
 * comp[v] receives the component index in emission order (sinks of the
 * condensation first); dead vertices get UINT32_MAX.  Returns #components.
 */
static int64_t tarjan_core(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                           const uint8_t *dead, uint32_t *comp)
{
    size_t n = (size_t)V + 1;
    uint32_t *idx = (uint32_t *)calloc(n, sizeof(uint32_t));
    uint32_t *low = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *stk = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *cs = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *ci = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint8_t *on = (uint8_t *)calloc(n, 1);
    uint32_t counter = 0, sp = 0;
    int64_t ncomp = 0;
    for (size_t v = 0; v < n; v++) comp[v] = UINT32_MAX;
    for (uint32_t s = 1; s <= V; s++) {
        if (idx[s] || (dead && dead[s])) continue;
        uint32_t csp = 0;
        idx[s] = low[s] = ++counter;
        stk[sp++] = s; on[s] = 1;
        cs[csp] = s; ci[csp] = Fr[s]; csp++;
        while (csp) {
            uint32_t v = cs[csp - 1];
            if (ci[csp - 1] < Fr[v + 1]) {
                uint32_t w = Fc[ci[csp - 1]++];
                if (dead && dead[w]) continue;
                if (!idx[w]) {
                    idx[w] = low[w] = ++counter;
                    stk[sp++] = w; on[w] = 1;
                    cs[csp] = w; ci[csp] = Fr[w]; csp++;
                } else if (on[w] && idx[w] < low[v]) {
                    low[v] = idx[w];
                }
            } else {
                if (low[v] == idx[v]) {
                    uint32_t x;
                    do {
                        x = stk[--sp];
                        on[x] = 0;
                        comp[x] = (uint32_t)ncomp;
                    } while (x != v);
                    ncomp++;
                }
                csp--;
                if (csp) {
                    uint32_t u = cs[csp - 1];
                    if (low[v] < low[u]) low[u] = low[v];
                }
            }
        }
    }
    free(idx); free(low); free(stk); free(cs); free(ci); free(on);
    return ncomp;
}

/* canonical labels from component ids: min vertex id per component */
static void comp_to_labels(uint32_t V, const uint32_t *comp, int64_t ncomp, uint32_t *lab)
{
    uint32_t *mn = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(ncomp ? ncomp : 1));
    for (int64_t c = 0; c < ncomp; c++) mn[c] = UINT32_MAX;
    for (uint32_t v = 1; v <= V; v++)
        if (comp[v] != UINT32_MAX && v < mn[comp[v]]) mn[comp[v]] = v;
    lab[0] = 0;
    for (uint32_t v = 1; v <= V; v++)
        lab[v] = comp[v] == UINT32_MAX ? 0 : mn[comp[v]];
    free(mn);
}

int64_t orc_tarjan(uint32_t V, const uint32_t *Fr, const uint32_t *Fc, uint32_t *lab)
{
    uint32_t *comp = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)V + 1));
    int64_t n = tarjan_core(V, Fr, Fc, NULL, comp);
    comp_to_labels(V, comp, n, lab);
    free(comp);
    return n;
}

/*
 * Kosaraju: iterative DFS finish order on the forward graph, then DFS in
 * decreasing finish order on the reverse graph.  Independent of Tarjan.
 */
int64_t orc_kosaraju(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                     const uint32_t *Br, const uint32_t *Bc, uint32_t *lab)
{
    size_t n = (size_t)V + 1;
    uint8_t *seen = (uint8_t *)calloc(n, 1);
    uint32_t *order = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *cs = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *ci = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t no = 0;
    for (uint32_t s = 1; s <= V; s++) {
        if (seen[s]) continue;
        uint32_t csp = 0;
        seen[s] = 1; cs[csp] = s; ci[csp] = Fr[s]; csp++;
        while (csp) {
            uint32_t v = cs[csp - 1];
            if (ci[csp - 1] < Fr[v + 1]) {
                uint32_t w = Fc[ci[csp - 1]++];
                if (!seen[w]) { seen[w] = 1; cs[csp] = w; ci[csp] = Fr[w]; csp++; }
            } else {
                order[no++] = v;
                csp--;
            }
        }
    }
    uint32_t *comp = (uint32_t *)malloc(n * sizeof(uint32_t));
    for (size_t v = 0; v < n; v++) comp[v] = UINT32_MAX;
    int64_t nc = 0;
    for (uint32_t k = no; k-- > 0;) {
        uint32_t s = order[k];
        if (comp[s] != UINT32_MAX) continue;
        uint32_t sp = 0;
        cs[sp++] = s; comp[s] = (uint32_t)nc;
        while (sp) {
            uint32_t v = cs[--sp];
            for (uint32_t i = Br[v]; i < Br[v + 1]; i++) {
                uint32_t w = Bc[i];
                if (comp[w] == UINT32_MAX) { comp[w] = (uint32_t)nc; cs[sp++] = w; }
            }
        }
        nc++;
    }
    comp_to_labels(V, comp, nc, lab);
    free(seen); free(order); free(cs); free(ci); free(comp);
    return nc;
}

/* Transitive closure with 64-bit rows; V <= 63.  u~v iff reach both ways. */
int64_t orc_brute(uint32_t V, const uint32_t *Fr, const uint32_t *Fc, uint32_t *lab)
{
    if (V > 63) return -1;
    uint64_t R[64] = { 0 };
    for (uint32_t u = 1; u <= V; u++) {
        R[u] |= 1ull << u;
        for (uint32_t i = Fr[u]; i < Fr[u + 1]; i++) R[u] |= 1ull << Fc[i];
    }
    for (uint32_t k = 1; k <= V; k++)          /* Warshall */
        for (uint32_t u = 1; u <= V; u++)
            if (R[u] >> k & 1) R[u] |= R[k];
    int64_t n = 0;
    lab[0] = 0;
    for (uint32_t v = 1; v <= V; v++) {
        uint32_t m = v;
        for (uint32_t u = 1; u < v; u++)
            if ((R[u] >> v & 1) && (R[v] >> u & 1)) { m = u; break; }
        lab[v] = m;
        if (m == v) n++;
    }
    return n;
}

void orc_scc_sizes(uint32_t V, const uint32_t *lab, uint32_t *sizes)
{
    memset(sizes, 0, sizeof(uint32_t) * ((size_t)V + 1));
    for (uint32_t v = 1; v <= V; v++) sizes[lab[v]]++;
}

/* ----------------------------------------------- phase restatements --- */

static int has_live_nbr(const uint32_t *R, const uint32_t *C, uint32_t v, const uint8_t *dead)
{
    for (uint32_t i = R[v]; i < R[v + 1]; i++)
        if (!dead[C[i]]) return 1;
    return 0;
}

/*
 * trim-1 fixpoint, one class (scc_kernels.cu:187-233; loop scc_WCC.cu:680-687):
 * a live vertex with no live in-neighbour or no live out-neighbour is
 * removed (a self-loop counts as a live neighbour, as in the reference).
 * The fixpoint is unique, so it is computed here by sequential peeling.
 */
int64_t orc_trim1_fixpoint(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                           const uint32_t *Br, const uint32_t *Bc, const uint8_t *dead,
                           uint8_t *removed)
{
    if (dead) memcpy(removed, dead, (size_t)V + 1);
    else memset(removed, 0, (size_t)V + 1);
    removed[0] = 1;
    int64_t total = 0;
    int changed = 1;
    while (changed) {
        changed = 0;
        for (uint32_t v = 1; v <= V; v++) {
            if (removed[v]) continue;
            if (!has_live_nbr(Br, Bc, v, removed) || !has_live_nbr(Fr, Fc, v, removed)) {
                removed[v] = 1;
                total++;
                changed = 1;
            }
        }
    }
    removed[0] = 0;
    if (dead)
        for (uint32_t v = 1; v <= V; v++) removed[v] = removed[v] && !dead[v];
    return total;
}

/* count live in-neighbours (capped at 2) and return the last one in *k */
static int live_deg2(const uint32_t *R, const uint32_t *C, uint32_t v, const uint8_t *dead,
                     uint32_t *k)
{
    int d = 0;
    for (uint32_t i = R[v]; i < R[v + 1]; i++) {
        uint32_t w = C[i];
        if (dead[w]) continue;
        if (++d == 2) return 2;
        *k = w;
    }
    return d;
}

static int has_arc(const uint32_t *R, const uint32_t *C, uint32_t v, uint32_t k)
{
    for (uint32_t i = R[v]; i < R[v + 1]; i++)
        if (C[i] == k) return 1;
    return 0;
}

/*
 * trim-2 on a snapshot (scc_kernels.cu:236-357): v is paired with k when
 * v's only live in-neighbour is k, v->k exists and k's only live
 * in-neighbour is v (case 1), or the mirror with out-neighbours (case 2).
 * pair_of[v] = k for each removed v (0 otherwise).  Returns #removed
 * vertices.  Every pair found here is also found by any concurrent pass
 * that starts from the same state.
 */
int64_t orc_trim2_snapshot(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                           const uint32_t *Br, const uint32_t *Bc,
                           const uint8_t *dead, uint32_t *pair_of)
{
    int64_t n = 0;
    memset(pair_of, 0, sizeof(uint32_t) * ((size_t)V + 1));
    for (uint32_t v = 1; v <= V; v++) {
        if (dead[v]) continue;
        uint32_t k = 0, kk = 0;
        int hit = 0;
        if (live_deg2(Br, Bc, v, dead, &k) == 1 && has_arc(Fr, Fc, v, k)
            && live_deg2(Br, Bc, k, dead, &kk) == 1)
            hit = 1;
        if (!hit) {
            k = 0;
            if (live_deg2(Fr, Fc, v, dead, &k) == 1 && has_arc(Br, Bc, v, k)
                && live_deg2(Fr, Fc, k, dead, &kk) == 1)
                hit = 1;
        }
        if (hit) { pair_of[v] = k; n++; }
    }
    return n;
}

/*
 * First-pivot rule (scc_kernels.cu:626-656) made deterministic: argmax over
 * live v of min(outdeg*indeg, 2^32-1) with global degrees, ties to the
 * smaller id.  Returns 0 when no vertex is live.
 */
uint32_t orc_first_pivot(uint32_t V, const uint32_t *Fr, const uint32_t *Br, const uint8_t *dead)
{
    uint32_t best = 0;
    uint64_t bk = 0;
    for (uint32_t v = 1; v <= V; v++) {
        if (dead && dead[v]) continue;
        uint64_t p = (uint64_t)(Fr[v + 1] - Fr[v]) * (uint64_t)(Br[v + 1] - Br[v]);
        if (p > 0xFFFFFFFFull) p = 0xFFFFFFFFull;
        uint64_t key = (p << 32) | (uint64_t)(0xFFFFFFFFu - v);
        if (!best || key > bk) { bk = key; best = v; }
    }
    return best;
}

/* reachable set from src over (R,C) within live vertices; out[v]=1 if reached */
int64_t orc_reach(uint32_t V, const uint32_t *R, const uint32_t *C, uint32_t src,
                  const uint8_t *dead, uint8_t *out)
{
    memset(out, 0, (size_t)V + 1);
    if (src < 1 || src > V || (dead && dead[src])) return 0;
    uint32_t *q = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)V + 1));
    uint32_t h = 0, t = 0;
    q[t++] = src; out[src] = 1;
    while (h < t) {
        uint32_t v = q[h++];
        for (uint32_t i = R[v]; i < R[v + 1]; i++) {
            uint32_t w = C[i];
            if (out[w] || (dead && dead[w])) continue;
            out[w] = 1;
            q[t++] = w;
        }
    }
    free(q);
    return t;
}

/*
 * Closed forms of the label-propagation fixpoints over the live subgraph:
 * min (resp. max) vertex id reachable from v.  Computed on the condensation
 * (Tarjan emits sinks first, so successors are final when a component is
 * processed).  Dead vertices get 0.
 */
static void reach_extreme(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                          const uint8_t *dead, uint32_t *out, int want_max)
{
    size_t n = (size_t)V + 1;
    uint32_t *comp = (uint32_t *)malloc(n * sizeof(uint32_t));
    int64_t nc = tarjan_core(V, Fr, Fc, dead, comp);
    uint32_t *best = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(nc ? nc : 1));
    uint32_t *start = (uint32_t *)calloc((size_t)nc + 1, sizeof(uint32_t));
    uint32_t *memb = (uint32_t *)malloc(n * sizeof(uint32_t));
    for (uint32_t v = 1; v <= V; v++)
        if (comp[v] != UINT32_MAX) start[comp[v] + 1]++;
    for (int64_t c = 0; c < nc; c++) start[c + 1] += start[c];
    uint32_t *fill = (uint32_t *)calloc((size_t)nc + 1, sizeof(uint32_t));
    for (uint32_t v = 1; v <= V; v++)
        if (comp[v] != UINT32_MAX) memb[start[comp[v]] + fill[comp[v]]++] = v;
    for (int64_t c = 0; c < nc; c++) {
        uint32_t b = want_max ? 0 : UINT32_MAX;
        for (uint32_t j = start[c]; j < start[c + 1]; j++) {
            uint32_t v = memb[j];
            if (want_max ? v > b : v < b) b = v;
            for (uint32_t i = Fr[v]; i < Fr[v + 1]; i++) {
                uint32_t w = Fc[i];
                if (comp[w] == UINT32_MAX || comp[w] == (uint32_t)c) continue;
                uint32_t x = best[comp[w]];
                if (want_max ? x > b : x < b) b = x;
            }
        }
        best[c] = b;
    }
    for (uint32_t v = 0; v <= V; v++)
        out[v] = (v == 0 || comp[v] == UINT32_MAX) ? 0 : best[comp[v]];
    free(comp); free(best); free(start); free(memb); free(fill);
}

void orc_min_reach_label(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                         const uint8_t *dead, uint32_t *out)
{
    reach_extreme(V, Fr, Fc, dead, out, 0);
}

void orc_max_reach_color(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                         const uint8_t *dead, uint32_t *out)
{
    reach_extreme(V, Fr, Fc, dead, out, 1);
}

/*
 * Local SCCs (the fixpoint of detectSCC's local loop, scc_WCC.cu:1175-1269,
 * whose searches are restricted to Pr[w]==Pr[v]): SCCs of the graph that
 * keeps only arcs with both ends in the same partition.
 */
int64_t orc_local_scc(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                      const uint32_t *part, uint32_t *lab)
{
    uint32_t *R = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)V + 2));
    uint32_t E = Fr[V + 1];
    uint32_t *C = (uint32_t *)malloc(sizeof(uint32_t) * (E ? E : 1));
    uint32_t m = 0;
    R[0] = R[1] = 0;
    for (uint32_t v = 1; v <= V; v++) {
        for (uint32_t i = Fr[v]; i < Fr[v + 1]; i++)
            if (part[Fc[i]] == part[v]) C[m++] = Fc[i];
        R[v + 1] = m;
    }
    int64_t n = orc_tarjan(V, R, C, lab);
    free(R);
    free(C);
    return n;
}
