/*
 * oracle_port.c -- multithreaded CPU restatement of the reference drivers.
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker's second opinion and
 * the CPU baseline reported beside the GPU numbers.
 *
 * vHong phase order (scc_WCC.cu:573-998):
 *   trim-1 fixpoint (680-687) -> first pivot (701-703) -> FW || BW from it
 *   (717-743) -> update (758) -> trim-1* (773-780) -> trim-2 (792-793) ->
 *   trim-1* (808-815) -> "WCC" min-label propagation with pointer jumping
 *   (830-840) -> loop { per-class pivots (860-862); FW || BW (876-902);
 *   update, exit when nothing is live (917-921) }.
 * vSlota phase order (scc_Coloring.cu:403-701):
 *   one trim-1 pass (507-508) -> first pivot FW || BW -> update ->
 *   loop { colour = max reachable id (596-602); roots colour==id (616);
 *   FW from roots inside the colour, each visited vertex joins the root's
 *   SCC (618-622); reset; exit when nothing is live (624-628) }.
 *
 * State: cls[v] = class id of a live vertex, DEAD once its SCC is known;
 * lab[v] = SCC representative; vis[v] bit0 = FW visited, bit1 = BW visited.
 * Class ids after a split are 3*pivot + (fw + 2*bw) (the reference's
 * 3*range + k of scc_kernels.cu:138, keyed by the pivot so ids never wrap).
 * Pivot per class = argmax of the 64-bit key (degree product, then smaller
 * id), so every class elects exactly one pivot.
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define DEAD 0xFFFFFFFFu
#define FV 1u
#define BV 2u

int orc_max_threads(void) { return omp_get_max_threads(); }

#define LD(p) __atomic_load_n((p), __ATOMIC_RELAXED)
#define ST(p, x) __atomic_store_n((p), (x), __ATOMIC_RELAXED)

typedef struct {
    uint32_t V;
    const uint32_t *Fr, *Fc, *Br, *Bc;
    uint32_t *cls, *lab, *aux;
    uint8_t *vis;
    uint64_t *piv;              /* 3(V+1)+3 keys, indexed by class id */
    uint32_t *fq[2], *bq[2];
    uint32_t **tbuf;            /* per-thread push buffers */
    size_t *tcap;
    int nth;
} port_t;

static uint64_t pivot_key(const port_t *P, uint32_t v)
{
    uint64_t p = (uint64_t)(P->Fr[v + 1] - P->Fr[v]) * (uint64_t)(P->Br[v + 1] - P->Br[v]);
    if (p > 0xFFFFFFFFull) p = 0xFFFFFFFFull;
    return (p << 32) | (uint64_t)(0xFFFFFFFFu - v);
}

static uint32_t key_vertex(uint64_t k) { return 0xFFFFFFFFu - (uint32_t)k; }

static void atomic_max64(uint64_t *a, uint64_t x)
{
    uint64_t o = LD(a);
    while (o < x && !__atomic_compare_exchange_n(a, &o, x, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
        ;
}

/* any live neighbour of v in class c over (R,C) (scc_kernels.cu:197-208) */
static int live_same(const port_t *P, const uint32_t *R, const uint32_t *C, uint32_t v, uint32_t c)
{
    for (uint32_t i = R[v]; i < R[v + 1]; i++)
        if (LD(&P->cls[C[i]]) == c) return 1;
    return 0;
}

/* one trim-1 sweep (scc_kernels.cu:187-233); returns #removed */
static int64_t trim1_pass(port_t *P)
{
    int64_t n = 0;
#pragma omp parallel for schedule(dynamic, 2048) reduction(+ : n) num_threads(P->nth)
    for (uint32_t v = 1; v <= P->V; v++) {
        uint32_t c = LD(&P->cls[v]);
        if (c == DEAD) continue;
        if (!live_same(P, P->Br, P->Bc, v, c) || !live_same(P, P->Fr, P->Fc, v, c)) {
            P->lab[v] = v;
            ST(&P->cls[v], DEAD);
            n++;
        }
    }
    return n;
}

static int64_t trim1_fix(port_t *P)
{
    int64_t tot = 0, n;
    do { n = trim1_pass(P); tot += n; } while (n);
    return tot;
}

/* live same-class degree capped at 2, last neighbour in *k */
static int deg2(const port_t *P, const uint32_t *R, const uint32_t *C, uint32_t v,
                uint32_t c, uint32_t *k)
{
    int d = 0;
    for (uint32_t i = R[v]; i < R[v + 1]; i++) {
        uint32_t w = C[i];
        if (LD(&P->cls[w]) != c) continue;
        if (++d == 2) return 2;
        *k = w;
    }
    return d;
}

static int arc_in(const uint32_t *R, const uint32_t *C, uint32_t v, uint32_t k)
{
    for (uint32_t i = R[v]; i < R[v + 1]; i++)
        if (C[i] == k) return 1;
    return 0;
}

/* one trim-2 pass (scc_kernels.cu:236-357); returns #pairs claimed */
static int64_t trim2_pass(port_t *P)
{
    int64_t n = 0;
#pragma omp parallel for schedule(dynamic, 2048) reduction(+ : n) num_threads(P->nth)
    for (uint32_t v = 1; v <= P->V; v++) {
        uint32_t c = LD(&P->cls[v]);
        if (c == DEAD) continue;
        uint32_t k = 0, kk;
        int hit = deg2(P, P->Br, P->Bc, v, c, &k) == 1 && arc_in(P->Fr, P->Fc, v, k)
                  && deg2(P, P->Br, P->Bc, k, c, &kk) == 1;
        if (!hit) {
            k = 0;
            hit = deg2(P, P->Fr, P->Fc, v, c, &k) == 1 && arc_in(P->Br, P->Bc, v, k)
                  && deg2(P, P->Fr, P->Fc, k, c, &kk) == 1;
        }
        if (!hit) continue;
        uint32_t lo = v < k ? v : k, hi = v < k ? k : v;
        uint32_t exp = c;
        /* the claim on the smaller id decides which thread records the pair */
        if (__atomic_compare_exchange_n(&P->cls[lo], &exp, DEAD, 0, __ATOMIC_RELAXED,
                                        __ATOMIC_RELAXED)) {
            P->lab[lo] = lo;
            P->lab[hi] = lo;
            ST(&P->cls[hi], DEAD);
            n++;
        }
    }
    return n;
}

/* push helpers: per-thread growable buffers, concatenated after a level */
static void tpush(port_t *P, int t, size_t *cnt, uint32_t v)
{
    if (cnt[t] == P->tcap[t]) {
        P->tcap[t] = P->tcap[t] ? 2 * P->tcap[t] : 4096;
        P->tbuf[t] = (uint32_t *)realloc(P->tbuf[t], P->tcap[t] * sizeof(uint32_t));
    }
    P->tbuf[t][cnt[t]++] = v;
}

static uint32_t gather(port_t *P, size_t *cnt, uint32_t *dst)
{
    size_t off = 0;
    for (int t = 0; t < P->nth; t++) {
        memcpy(dst + off, P->tbuf[t], cnt[t] * sizeof(uint32_t));
        off += cnt[t];
    }
    return (uint32_t)off;
}

/*
 * One direction, one BFS level within each vertex's class (fwd/bwd,
 * scc_kernels.cu:360-445, data-driven instead of topology-driven).
 */
static uint32_t expand(port_t *P, const uint32_t *R, const uint32_t *C, uint8_t bit,
                       const uint32_t *in, uint32_t nin, uint32_t *out)
{
    size_t cnt[256] = { 0 };
#pragma omp parallel num_threads(P->nth)
    {
        int t = omp_get_thread_num();
#pragma omp for schedule(dynamic, 64)
        for (uint32_t j = 0; j < nin; j++) {
            uint32_t v = in[j];
            uint32_t c = LD(&P->cls[v]);
            for (uint32_t i = R[v]; i < R[v + 1]; i++) {
                uint32_t w = C[i];
                if (LD(&P->cls[w]) != c || (LD(&P->vis[w]) & bit)) continue;
                if (!(__atomic_fetch_or(&P->vis[w], bit, __ATOMIC_RELAXED) & bit))
                    tpush(P, t, cnt, w);
            }
        }
    }
    return gather(P, cnt, out);
}

/* FW || BW from the seeded queues until both are exhausted */
static void fwbw(port_t *P, uint32_t nf, uint32_t nb, int do_bw)
{
    int cf = 0, cb = 0;
    while (nf || (do_bw && nb)) {
        if (nf) { nf = expand(P, P->Fr, P->Fc, FV, P->fq[cf], nf, P->fq[cf ^ 1]); cf ^= 1; }
        if (do_bw && nb) { nb = expand(P, P->Br, P->Bc, BV, P->bq[cb], nb, P->bq[cb ^ 1]); cb ^= 1; }
    }
}

/*
 * Elect one pivot per live class (pollForPivots/selectPivots,
 * scc_kernels.cu:4-19, 57-74 with the deterministic key) and seed the queues.
 */
static uint32_t elect_and_seed(port_t *P, int seed_bw)
{
    const uint32_t V = P->V;
#pragma omp parallel for schedule(static) num_threads(P->nth)
    for (uint32_t v = 1; v <= V; v++) {
        uint32_t c = P->cls[v];
        if (c != DEAD) ST(&P->piv[c], 0);
    }
#pragma omp parallel for schedule(static) num_threads(P->nth)
    for (uint32_t v = 1; v <= V; v++) {
        uint32_t c = P->cls[v];
        if (c != DEAD) atomic_max64(&P->piv[c], pivot_key(P, v));
    }
    uint32_t n = 0;
    for (uint32_t v = 1; v <= V; v++) {   /* sequential: keeps queues ordered */
        P->vis[v] = 0;
        uint32_t c = P->cls[v];
        if (c != DEAD && key_vertex(P->piv[c]) == v) {
            P->vis[v] = FV | (seed_bw ? BV : 0);
            P->fq[0][n] = v;
            P->bq[0][n] = v;
            n++;
        }
    }
    return n;
}

/* update (scc_kernels.cu:124-142): FW&BW -> SCC of the pivot; else split */
static int64_t update(port_t *P, int64_t *live)
{
    int64_t found = 0, lv = 0;
#pragma omp parallel for schedule(static) reduction(+ : found, lv) num_threads(P->nth)
    for (uint32_t v = 1; v <= P->V; v++) {
        uint32_t c = P->cls[v];
        if (c == DEAD) continue;
        uint32_t p = key_vertex(P->piv[c]);
        uint8_t f = P->vis[v];
        if ((f & FV) && (f & BV)) {
            P->lab[v] = p;
            P->cls[v] = DEAD;
            found += (v == p);
        } else {
            P->cls[v] = 3u * p + (f & FV ? 1u : 0u) + (f & BV ? 2u : 0u);
            lv++;
        }
    }
    *live = lv;
    return found;
}

/* "WCC": min reachable id by propagateRange1 + propagateRange2 (scc_kernels.cu:496-536) */
static void min_label(port_t *P)
{
    const uint32_t V = P->V;
#pragma omp parallel for schedule(static) num_threads(P->nth)
    for (uint32_t v = 1; v <= V; v++)
        if (P->cls[v] != DEAD) P->cls[v] = v;
    int changed = 1;
    while (changed) {
        changed = 0;
#pragma omp parallel for schedule(dynamic, 2048) reduction(| : changed) num_threads(P->nth)
        for (uint32_t v = 1; v <= V; v++) {
            uint32_t m0 = LD(&P->cls[v]);
            if (m0 == DEAD) continue;
            uint32_t m = m0;
            for (uint32_t i = P->Fr[v]; i < P->Fr[v + 1]; i++) {
                uint32_t x = LD(&P->cls[P->Fc[i]]);
                if (x < m) m = x;           /* DEAD is the largest value */
            }
            for (uint32_t x; (x = LD(&P->cls[m])) < m;) m = x;
            if (m < m0) { ST(&P->cls[v], m); changed = 1; }
        }
    }
#pragma omp parallel for schedule(static) num_threads(P->nth)
    for (uint32_t v = 1; v <= V; v++)
        if (P->cls[v] != DEAD) P->cls[v] *= 3u;
}

/* colour = max reachable id (colorPropagation, scc_kernels.cu:539-564) */
static void max_color(port_t *P)
{
    const uint32_t V = P->V;
    uint32_t *a = P->aux;
#pragma omp parallel for schedule(static) num_threads(P->nth)
    for (uint32_t v = 0; v <= V; v++) a[v] = (v && P->cls[v] != DEAD) ? v : 0;
    int changed = 1;
    while (changed) {
        changed = 0;
#pragma omp parallel for schedule(dynamic, 2048) reduction(| : changed) num_threads(P->nth)
        for (uint32_t v = 1; v <= V; v++) {
            uint32_t m0 = LD(&a[v]);
            if (!m0) continue;
            uint32_t m = m0;
            for (uint32_t i = P->Fr[v]; i < P->Fr[v + 1]; i++) {
                uint32_t x = LD(&a[P->Fc[i]]);
                if (x > m) m = x;           /* dead vertices hold 0 */
            }
            for (uint32_t x; (x = LD(&a[m])) > m;) m = x;
            if (m > m0) { ST(&a[v], m); changed = 1; }
        }
    }
}

static void canonical(port_t *P)
{
    const uint32_t V = P->V;
    uint32_t *mn = P->aux;
    for (uint32_t v = 0; v <= V; v++) mn[v] = DEAD;
    for (uint32_t v = 1; v <= V; v++)
        if (v < mn[P->lab[v]]) mn[P->lab[v]] = v;
#pragma omp parallel for schedule(static) num_threads(P->nth)
    for (uint32_t v = 1; v <= V; v++) P->lab[v] = mn[P->lab[v]];
    P->lab[0] = 0;
}

static int port_init(port_t *P, uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                     const uint32_t *Br, const uint32_t *Bc, int nth, uint32_t *lab)
{
    memset(P, 0, sizeof(*P));
    P->V = V; P->Fr = Fr; P->Fc = Fc; P->Br = Br; P->Bc = Bc;
    P->nth = nth > 0 ? nth : omp_get_max_threads();
    if (P->nth > 256) P->nth = 256;
    size_t n = (size_t)V + 1;
    P->lab = lab;
    P->cls = (uint32_t *)malloc(n * 4);
    P->aux = (uint32_t *)malloc(n * 4);
    P->vis = (uint8_t *)calloc(n, 1);
    P->piv = (uint64_t *)calloc(3 * n + 3, 8);
    for (int i = 0; i < 2; i++) {
        P->fq[i] = (uint32_t *)malloc(n * 4);
        P->bq[i] = (uint32_t *)malloc(n * 4);
    }
    P->tbuf = (uint32_t **)calloc((size_t)P->nth, sizeof(uint32_t *));
    P->tcap = (size_t *)calloc((size_t)P->nth, sizeof(size_t));
    P->cls[0] = DEAD;
    lab[0] = 0;
    for (uint32_t v = 1; v <= V; v++) { P->cls[v] = 0; lab[v] = v; }
    return 0;
}

static void port_fini(port_t *P)
{
    free(P->cls); free(P->aux); free(P->vis); free(P->piv);
    for (int i = 0; i < 2; i++) { free(P->fq[i]); free(P->bq[i]); }
    for (int t = 0; t < P->nth; t++) free(P->tbuf[t]);
    free(P->tbuf); free(P->tcap);
}

static int64_t count_roots(const port_t *P)
{
    int64_t n = 0;
    for (uint32_t v = 1; v <= P->V; v++) n += P->lab[v] == v;
    return n;
}

int64_t orc_vhong(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                  const uint32_t *Br, const uint32_t *Bc, int t1, int t2,
                  int nthreads, uint32_t *lab, orc_stats *st)
{
    port_t P;
    orc_stats s;
    memset(&s, 0, sizeof(s));
    port_init(&P, V, Fr, Fc, Br, Bc, nthreads, lab);
    int64_t live = V;
    if (t1) s.num_trim1 += trim1_fix(&P);
    /* first pivot: one class, so elect_and_seed elects the global argmax */
    uint32_t np = elect_and_seed(&P, 1);
    if (np) {
        fwbw(&P, np, np, 1);
        s.num_pivot += update(&P, &live);
    }
    if (t1) s.num_trim1 += trim1_fix(&P);
    if (t2) s.num_trim2 += trim2_pass(&P);
    if (t1) s.num_trim1 += trim1_fix(&P);
    min_label(&P);
    for (;;) {
        np = elect_and_seed(&P, 1);
        if (!np) break;
        s.rounds++;
        fwbw(&P, np, np, 1);
        s.num_pivot += update(&P, &live);
        if (!live) break;
    }
    canonical(&P);
    s.num_sccs = count_roots(&P);
    port_fini(&P);
    if (st) *st = s;
    return s.num_sccs;
}

int64_t orc_vslota(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                   const uint32_t *Br, const uint32_t *Bc, int t1,
                   int nthreads, uint32_t *lab, orc_stats *st)
{
    port_t P;
    orc_stats s;
    memset(&s, 0, sizeof(s));
    port_init(&P, V, Fr, Fc, Br, Bc, nthreads, lab);
    int64_t live = V;
    if (t1) s.num_trim1 += trim1_pass(&P);      /* one pass: scc_Coloring.cu:507-508 */
    uint32_t np = elect_and_seed(&P, 1);
    if (np) {
        fwbw(&P, np, np, 1);
        s.num_pivot += update(&P, &live);
    }
    for (;;) {
        int64_t lv = 0;
        for (uint32_t v = 1; v <= V; v++) lv += P.cls[v] != DEAD;
        if (!lv) break;
        s.rounds++;
        max_color(&P);
        /* roots (selectPivotColoring, scc_kernels.cu:567-580): class = colour */
        uint32_t n = 0;
        for (uint32_t v = 1; v <= V; v++) {
            P.vis[v] = 0;
            if (P.cls[v] == DEAD) continue;
            P.cls[v] = P.aux[v];
            if (P.aux[v] == v) { P.vis[v] = FV; P.fq[0][n++] = v; }
        }
        fwbw(&P, n, 0, 0);
        /* fwdColoring marks every expanded vertex finished (scc_kernels.cu:607) */
        for (uint32_t v = 1; v <= V; v++)
            if (P.cls[v] != DEAD && (P.vis[v] & FV)) {
                P.lab[v] = P.cls[v];
                P.cls[v] = DEAD;
            }
        s.num_pivot += n;
    }
    canonical(&P);
    s.num_sccs = count_roots(&P);
    port_fini(&P);
    if (st) *st = s;
    return s.num_sccs;
}
