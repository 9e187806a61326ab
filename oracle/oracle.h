/*
 * oracle.h -- CPU oracle for the SCC engine.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * `scc` executable, the Python package) links, loads or calls this code.
 * It is used by tests/, by __graft_entry__.smoke() and by the cpu_baseline
 * leg of bench.py, and only as the checker or the reported CPU baseline.
 *
 * What it restates (reference = Amit104/DetectingSCC):
 *   - orc_load_gtgraph / orc_build_csr : src/load.cpp:12-113 (GTgraph text ->
 *     1-based CSR + reverse CSR, dummy vertex 0, arcs sorted by (src,dst),
 *     in-lists ascending by source, duplicates and self-loops kept).
 *   - orc_load_partition : src/load.cpp:117-130 (one partition id per line).
 *   - orc_trim1_fixpoint : scc_kernels.cu:187-233 iterated as in
 *     scc_WCC.cu:680-687 (single class, every vertex live at the start).
 *   - orc_trim2_snapshot : scc_kernels.cu:236-357 evaluated on a snapshot.
 *   - orc_first_pivot    : scc_kernels.cu:626-656 with the deterministic
 *     key (outdeg*indeg saturated to 32 bits, ties to the smaller id).
 *   - orc_reach          : one direction of fwd/bwd (scc_kernels.cu:360-445)
 *     iterated to its fixpoint = the reachable set.
 *   - orc_min_reach_label: assignUniqueRange + propagateRange1/2 fixpoint
 *     (scc_kernels.cu:477-536, driven at scc_WCC.cu:830-840).
 *   - orc_max_reach_color: colorPropagation fixpoint (scc_kernels.cu:539-564,
 *     driven at scc_Coloring.cu:596-602).
 *   - orc_vhong / orc_vslota : the vHong (scc_WCC.cu:573-998) and vSlota
 *     (scc_Coloring.cu:403-701) phase orders, multithreaded with OpenMP.
 *     This is the reported CPU baseline ("port").
 *   - orc_local_scc      : SCCs of the union of partition-induced subgraphs
 *     (what detectSCC's local loop computes, scc_WCC.cu:1175-1269).
 *
 * Ground truth (pins the above): orc_tarjan (Tarjan 1972, iterative),
 * orc_kosaraju (Kosaraju/Sharir, iterative) and orc_brute (transitive
 * closure by bitsets, V <= 64).  Labels are canonical: lab[v] = the minimum
 * vertex id of v's SCC, lab[0] = 0 for the dummy vertex.
 *
 * Every array indexed by vertex has V+1 entries (index 0 is the dummy).
 * Fr/Br have V+2 entries, Fc/Bc have E entries (reference layout, A1).
 */
#ifndef SCC_ORACLE_H
#define SCC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int64_t num_sccs;
    int64_t num_trim1;   /* vertices removed by trim-1 */
    int64_t num_trim2;   /* pairs removed by trim-2 */
    int64_t num_pivot;   /* SCCs found by a pivot search */
    int64_t rounds;      /* outer rounds of the main loop */
} orc_stats;

/* graph I/O (load.cpp) ; return 0 on success, <0 on error */
int  orc_load_gtgraph(const char *path, uint32_t *V, uint32_t *E,
                      uint32_t **Fr, uint32_t **Fc, uint32_t **Br, uint32_t **Bc);
int  orc_load_partition(const char *path, uint32_t V, uint32_t *part /* V+2 */);
void orc_build_csr(uint32_t V, uint32_t E, const uint32_t *src, const uint32_t *dst,
                   uint32_t *Fr, uint32_t *Fc, uint32_t *Br, uint32_t *Bc);
void orc_free(void *p);

/* ground truth */
int64_t orc_tarjan(uint32_t V, const uint32_t *Fr, const uint32_t *Fc, uint32_t *lab);
int64_t orc_kosaraju(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                     const uint32_t *Br, const uint32_t *Bc, uint32_t *lab);
int64_t orc_brute(uint32_t V, const uint32_t *Fr, const uint32_t *Fc, uint32_t *lab);
/* sizes[r] = |SCC with label r| for roots r (lab[r]==r), 0 elsewhere */
void    orc_scc_sizes(uint32_t V, const uint32_t *lab, uint32_t *sizes);

/* phase restatements */
int64_t  orc_trim1_fixpoint(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                            const uint32_t *Br, const uint32_t *Bc, const uint8_t *dead,
                            uint8_t *removed);
int64_t  orc_trim2_snapshot(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                            const uint32_t *Br, const uint32_t *Bc,
                            const uint8_t *dead, uint32_t *pair_of);
uint32_t orc_first_pivot(uint32_t V, const uint32_t *Fr, const uint32_t *Br, const uint8_t *dead);
int64_t  orc_reach(uint32_t V, const uint32_t *R, const uint32_t *C, uint32_t src,
                   const uint8_t *dead, uint8_t *out);
void     orc_min_reach_label(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                             const uint8_t *dead, uint32_t *out);
void     orc_max_reach_color(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                             const uint8_t *dead, uint32_t *out);
int64_t  orc_local_scc(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                       const uint32_t *part, uint32_t *lab);

/* CPU ports of the reference drivers (OpenMP, nthreads<=0 = all cores) */
int64_t orc_vhong(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                  const uint32_t *Br, const uint32_t *Bc, int t1, int t2,
                  int nthreads, uint32_t *lab, orc_stats *st);
int64_t orc_vslota(uint32_t V, const uint32_t *Fr, const uint32_t *Fc,
                   const uint32_t *Br, const uint32_t *Bc, int t1,
                   int nthreads, uint32_t *lab, orc_stats *st);
int     orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
