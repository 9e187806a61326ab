/*
 * scc_amd.h -- C ABI of the MI355X strongly-connected-components engine.
 *
 * Reference: Amit104/DetectingSCC.  Each entry point below names the
 * reference interface it replaces (file:line).  Conventions:
 *   - every vertex-indexed array follows the reference layout (A1):
 *     vertices are 1..V, index 0 is a dummy vertex with no arcs,
 *     row arrays (Fr, Br) have V+2 entries, column arrays (Fc, Bc) E entries,
 *     Fc sorted by (src, dst), each in-list of Bc ascending by source;
 *   - SCC labels are canonical: lab[v] = the smallest vertex id in v's SCC,
 *     lab[0] = 0; arrays of labels have V+1 entries;
 *   - functions return scc_status (0 = success) and never exit or throw;
 *     scc_last_error() gives a message for the calling thread;
 *   - host buffers are owned by the caller unless documented otherwise;
 *     device memory is owned by the context; one context per host thread.
 */
#ifndef SCC_AMD_H
#define SCC_AMD_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif
#pragma GCC visibility push(default)

typedef enum {
    SCC_OK = 0,
    SCC_ERR_INVALID = 1,     /* bad argument */
    SCC_ERR_IO = 2,          /* file cannot be opened or read */
    SCC_ERR_PARSE = 3,       /* malformed GTgraph / partition file */
    SCC_ERR_NOMEM = 4,       /* host or device allocation failed */
    SCC_ERR_DEVICE = 5,      /* HIP runtime error */
    SCC_ERR_BUDGET = 6,      /* request larger than the HBM budget */
    SCC_ERR_UNSUPPORTED = 7, /* option combination not supported */
    SCC_ERR_NOTREADY = 8     /* nothing to return yet */
} scc_status;

const char *scc_status_string(int status);
const char *scc_last_error(void);
const char *scc_version(void);

/* ------------------------------------------------------------ graphs --- */

/* Host CSR + reverse CSR in the reference layout; arrays from malloc. */
typedef struct {
    uint32_t V;
    uint32_t E;
    uint32_t *Fr, *Fc, *Br, *Bc;
} scc_host_graph;

void scc_host_graph_free(scc_host_graph *g);

/*
 * GTgraph loader.  Replaces loadFullGraph (include/load.h:6,
 * src/load.cpp:12-113): skips 7 header lines, reads "p sp V E" and E arcs
 * "a u v w" (weight ignored), sorts arcs by (u, v), builds CSR and reverse
 * CSR.  Unlike the reference, ids outside [1, V], short files and bad
 * headers return SCC_ERR_PARSE, and no partition file is opened (the
 * hard-coded path of load.cpp:117-122 becomes scc_load_partition).
 */
scc_status scc_load_gtgraph(const char *path, scc_host_graph *out);

/* Partition vector (src/load.cpp:117-130): V ids, one per line, into
 * part[1..V]; part has V+2 entries. */
scc_status scc_load_partition(const char *path, uint32_t V, uint32_t *part);

/* Arc list -> reference CSR layout (src/load.cpp:47-113), multithreaded. */
scc_status scc_csr_from_arcs(uint32_t V, uint64_t E, const uint32_t *src, const uint32_t *dst,
                             scc_host_graph *out);

/* Writers for the on-disk formats: GTgraph text (the loader's input) and
 * the METIS .graph of src/MetisInput.cpp:115-131 (out-list then in-list). */
scc_status scc_write_gtgraph(const char *path, const scc_host_graph *g);
scc_status scc_write_metis(const char *path, const scc_host_graph *g);

/* Binary CSR cache (no reference counterpart; the reference re-parses and
 * re-sorts the GTgraph text on every run, load.cpp:32-113): magic
 * "SCCCSR01", V and E as u32, 8 reserved bytes, then Fr[V+2], Fc[E],
 * Br[V+2], Bc[E] as little-endian u32.  The loader checks the reference
 * layout (rows from 0,0 to E, monotone, rows sorted, ids in [1, V], the
 * reverse CSR's in-degrees equal to the forward columns' counts). */
scc_status scc_write_csr(const char *path, const scc_host_graph *g);
scc_status scc_load_csr(const char *path, scc_host_graph *out);
/* Either format, told apart by the binary CSR magic (the `scc -f` path). */
scc_status scc_load_graph(const char *path, scc_host_graph *out);

/*
 * Seeded synthetic graphs (no reference counterpart; the reference ships
 * only sample files).  Counter-based RNG: output does not depend on the
 * number of host threads.
 *   ER:    `draws` arcs (u, v) uniform in [1, V]^2, u == v dropped,
 *          duplicates kept, plus `cycles` directed cycles of length
 *          uniform in [cmin, cmax] over distinct vertices.
 *   R-MAT: V = 2^scale, E arcs, quadrant probabilities (a, b, c, 1-a-b-c),
 *          seeded vertex permutation, self-loops and duplicates kept.
 *   grid:  side x side, id = r*side + c + 1, each N/S/E/W arc inside the
 *          grid present with probability p.
 *   communities: V vertices in K equal planted communities (after a seeded
 *          permutation); arc count E = V*deg; a fraction `inter` of arcs
 *          join uniform random vertices, the rest stay in the source's
 *          community.  part[1..V] = community id (0-based), V+2 entries.
 */
scc_status scc_gen_er(uint32_t V, uint64_t draws, uint32_t cycles, uint32_t cmin, uint32_t cmax,
                      uint64_t seed, scc_host_graph *out);
scc_status scc_gen_rmat(uint32_t scale, uint64_t E, double a, double b, double c, uint64_t seed,
                        scc_host_graph *out);
scc_status scc_gen_grid(uint32_t side, double p, uint64_t seed, scc_host_graph *out);
/* The generators' seeded permutation of [0, n): out[i] for i < n (a
 * four-round Feistel network with cycle walking, so R-MAT and community
 * graphs can be permuted per arc on the host or the device alike). */
scc_status scc_seeded_permutation(uint32_t n, uint64_t seed, uint32_t *out);
scc_status scc_gen_communities(uint32_t V, uint32_t K, double deg, double inter, uint64_t seed,
                               scc_host_graph *out, uint32_t *part);

/* ------------------------------------------------------- device engine --- */

typedef struct scc_ctx scc_ctx;

/* Algorithms of the reference CLI (src/scc.cu:78-102). */
#define SCC_ALGO_VHONG  'g' /* vHong  scc_WCC.cu:573-998 */
#define SCC_ALGO_VSLOTA 'h' /* vSlota scc_Coloring.cu:403-701 */
#define SCC_ALGO_WHONG  'x' /* wHong  scc_WCC.cu:53-570 */
#define SCC_ALGO_WSLOTA 'y' /* wSlota scc_Coloring.cu:6-400 */
#define SCC_ALGO_DETECT 'd' /* detectSCC scc_WCC.cu:1000-1617 */

/* Frontier expansion strategy of the reachability kernels. */
#define SCC_EXPAND_AUTO   0 /* direction-optimising (SCC_EXPAND_DO) for every algorithm */
#define SCC_EXPAND_VERTEX 1 /* one lane per frontier vertex (fwd/bwd, scc_kernels.cu:360-445) */
#define SCC_EXPAND_VWARP  2 /* W lanes per vertex (fwd_warp/bwd_warp, scc_kernels.h:46-105) */
#define SCC_EXPAND_LB     3 /* workgroup scan-based gather + hub splitting */
#define SCC_EXPAND_DO     4 /* direction-optimising: LB top-down + bottom-up passes */

typedef struct {
    char algo;        /* SCC_ALGO_* */
    int trim1;        /* -p */
    int trim2;        /* -q */
    int warp;         /* -w: 1,2,4,8,16,32,64 */
    int expand;       /* SCC_EXPAND_* */
    int block;        /* reserved, 0 */
    int check_every;  /* BFS levels between host checks for the vertex, virtual-warp and LB
                         mappings, 0 = 2 (direction-optimised searches decide on the device) */
    int sizes;        /* 1: also compute stats.largest and the first search's
                         vertex/arc counts (two more sweeps) */
    int phase_times;  /* 1: fill stats.ms_phase (a device timestamp at every phase
                         boundary, each a few microseconds of idle GPU); 0: only
                         ms_scc and ms_first_reach are timed */
} scc_opts;

void scc_opts_default(scc_opts *o);

#define SCC_PHASE_TRIM1   0
#define SCC_PHASE_TRIM2   1
#define SCC_PHASE_PIVOT   2
#define SCC_PHASE_REACH   3
#define SCC_PHASE_UPDATE  4
#define SCC_PHASE_WCC     5
#define SCC_PHASE_COLOR   6
#define SCC_PHASE_LOCAL   7
#define SCC_PHASE_MERGE   8
#define SCC_PHASE_LABEL   9
#define SCC_PHASE_COUNT  10

typedef struct {
    uint32_t num_sccs;       /* "Number Of Sccs" (scc_WCC.cu:966) */
    uint32_t num_trim1;      /* "numberOf1Sccs" (scc_WCC.cu:948) */
    uint32_t num_trim2;      /* 2-SCCs removed by trim-2 */
    uint32_t num_pivot;      /* SCCs found by a pivot / colour-root search */
    uint32_t largest;        /* size of the largest SCC (0 unless opts.sizes) */
    uint32_t rounds;         /* outer rounds of the main loop */
    uint32_t levels;         /* BFS levels launched, both directions */
    uint32_t first_pivot;    /* vertex elected by the first-pivot rule */
    uint64_t first_fw_vertices, first_fw_arcs; /* |R|, |E_R| of the first FW (opts.sizes) */
    uint64_t first_bw_vertices, first_bw_arcs; /* |R|, |E_R| of the first BW (opts.sizes) */
    double ms_scc;           /* CSR resident -> labels resident (device clock) */
    double ms_first_reach;   /* first-pivot FW||BW (device clock) */
    double ms_phase[SCC_PHASE_COUNT];   /* with opts.phase_times, else 0 */
    /* Work units of the first search when it is direction-optimised and
     * opts.sizes is set (else 0), both directions: [0] vertices expanded
     * top-down, [1] arcs scanned top-down, [2] vertices claimed, [3]
     * vertices a bottom-up pass looked at, [4] unvisited candidates, [5]
     * first parents probed, [6] row-offset pairs read after a first-parent
     * miss, [7] further parents probed.  Their bytes are 12, 8, 4, 1/8, 4,
     * 8, 8 and 12 each (the search's algorithmic traffic). */
    uint64_t first_work[8];
} scc_stats;

/* Create a context on HIP device `device` (the reference hard-codes
 * cudaSetDevice(7), scc_WCC.cu:575; here it is a parameter). */
scc_status scc_ctx_create(int device, scc_ctx **out);
void scc_ctx_destroy(scc_ctx *ctx);

/* Copy a graph to the device and keep it resident (replaces the
 * per-driver cudaMalloc + H2D, scc_WCC.cu:608-628). */
scc_status scc_graph_upload(scc_ctx *ctx, uint32_t V, uint32_t E, const uint32_t *Fr,
                            const uint32_t *Fc, const uint32_t *Br, const uint32_t *Bc);
/* Partition vector for SCC_ALGO_DETECT (the Pr of scc_WCC.cu:1072). */
scc_status scc_partition_upload(scc_ctx *ctx, const uint32_t *part);

/* Run one SCC decomposition on the resident graph; labels stay resident. */
scc_status scc_run(scc_ctx *ctx, const scc_opts *opts, scc_stats *stats);
/* Copy the canonical labels of the last run (V+1 entries). */
scc_status scc_labels(scc_ctx *ctx, uint32_t *lab);
/* The condensation DAG of the last scc_run (no reference counterpart: the
 * reference prints only a count, scc_WCC.cu:966): node k (1-based) is the
 * SCC with the k-th smallest canonical label, one arc per distinct ordered
 * pair of SCCs joined by at least one arc, in the reference layout (arrays
 * from malloc, free with scc_host_graph_free).  node_of (V+1 entries, may be
 * NULL) receives each vertex's node. */
scc_status scc_condensation(scc_ctx *ctx, scc_host_graph *dag, uint32_t *node_of);

/* Upload + run + labels in one call. */
scc_status scc_solve(scc_ctx *ctx, const scc_host_graph *g, const uint32_t *part,
                     const scc_opts *opts, uint32_t *lab, scc_stats *stats);

/* SCC size histogram of the last scc_run, computed on the device from its
 * canonical labels: (size, count) pairs ascending by size, up to `cap` of
 * them written to pairs[2k], pairs[2k+1]; *npairs = the number of distinct
 * sizes (at most floor(sqrt(2V)) + 1).  *ms (may be NULL) = device time.
 * The reference prints only a count (scc_WCC.cu:966); this output has no
 * counterpart there. */
scc_status scc_histogram(scc_ctx *ctx, uint64_t *pairs, uint32_t cap, uint32_t *npairs, double *ms);

/* Allocation-free variant of scc_histogram.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same outputs as scc_histogram (pairs ascending by size, up
 * to `cap` of them; *npairs = distinct sizes), no timing.  It calls no
 * hipMalloc, hipFree or other allocating routine: its staging lives in a
 * static device array of 262,400 bytes per GPU (held until the process
 * ends), sized for at most 16,384 distinct SCC sizes, so graphs of more
 * than 134,225,920 vertices are rejected with SCC_ERR_INVALID.  Every call
 * on a GPU shares that array: calls that overlap on one GPU race. */
scc_status scc_histogram_noalloc(scc_ctx *ctx, uint64_t *pairs, uint32_t cap, uint32_t *npairs);

/* SCC size histogram from labels: pairs (size, count) ascending by size,
 * at most `cap` pairs; *npairs receives the number of distinct sizes. */
scc_status scc_size_histogram(uint32_t V, const uint32_t *lab, uint64_t *pairs, uint32_t cap,
                              uint32_t *npairs);

/* ------------------------------------------------ phase entry points --- */
/*
 * Single phases on the resident graph, for kernel-level checks.  `dead`
 * (V+1 bytes, may be NULL) marks vertices already finished; all other
 * vertices form one class.
 */
/* trim-1 fixpoint (scc_kernels.cu:187-233 iterated); removed[v]=1 */
scc_status scc_phase_trim1(scc_ctx *ctx, const uint8_t *dead, uint8_t *removed);
/* one trim-2 pass (scc_kernels.cu:236-357); pair[v] = min id of v's pair, 0 if kept */
scc_status scc_phase_trim2(scc_ctx *ctx, const uint8_t *dead, uint32_t *pair);
/* first-pivot rule (scc_kernels.cu:626-656, deterministic key) */
scc_status scc_phase_first_pivot(scc_ctx *ctx, const uint8_t *dead, uint32_t *pivot);
/* reachability from src; dir 0 = FW over Fc, 1 = BW over Bc (scc_kernels.cu:360-445,
 * scc_kernels.h:46-105); out[v]=1 if reached; *levels = BFS levels */
scc_status scc_phase_reach(scc_ctx *ctx, uint32_t src, int dir, int expand, int warp,
                           const uint8_t *dead, uint8_t *out, uint32_t *levels, double *ms);
/* min reachable id (propagateRange1/2, scc_kernels.cu:496-536); 0 for dead */
scc_status scc_phase_min_label(scc_ctx *ctx, const uint8_t *dead, uint32_t *out);
/* max reachable id (colorPropagation, scc_kernels.cu:539-564); 0 for dead */
scc_status scc_phase_max_color(scc_ctx *ctx, const uint8_t *dead, uint32_t *out);
/* local SCCs of detectSCC (scc_WCC.cu:1122-1269): global trim-1, then SCCs
 * of the union of partition-induced subgraphs; canonical labels (V+1) of the
 * uploaded graph and partition (trim-2 is left out: a 2-cycle across two
 * partitions is one global SCC but two local ones) */
scc_status scc_phase_local(scc_ctx *ctx, uint32_t *out);

/* ----------------------------------------------------------- batches --- */
/*
 * Many independent graphs solved together (no reference counterpart: the
 * reference solves one graph per process).  Host layout: graph i's rows
 * Fr/Br (V[i]+2 entries, reference layout, local ids 1..V[i]) start at
 * row_off[i]; its columns Fc/Bc (E[i] entries) start at col_off[i].
 */
typedef struct {
    uint32_t num_graphs;
    uint32_t *V, *E;
    uint64_t *row_off, *col_off;  /* num_graphs+1 entries each */
    uint32_t *Fr, *Fc, *Br, *Bc;
    uint8_t *kind;                /* generator kind (0 ER, 1 R-MAT, 2 planted) or NULL */
} scc_batch_host;

void scc_batch_host_free(scc_batch_host *b);

/* The batched mix: V log-uniform in [2^lgmin, 2^lgmax], mean degree uniform
 * in [dmin, dmax], 40% ER / 40% R-MAT / 20% planted cycles (2..512) + ER. */
scc_status scc_gen_batch_mix(uint32_t num_graphs, uint32_t lgmin, uint32_t lgmax, double dmin,
                             double dmax, uint64_t seed, scc_batch_host *out);
/* The serving stream: V lognormal with the given median and sigma, clamped
 * to [2, vcap], mean out-degree `degree`, 50% ER / 50% R-MAT. */
scc_status scc_gen_serving_mix(uint32_t num_graphs, double median_v, double sigma, uint32_t vcap,
                               double degree, uint64_t seed, scc_batch_host *out);

typedef struct {
    char algo;       /* 'g' Hong phase order or 'h' Slota colouring */
    int trim1, trim2;
    int lanes;       /* lanes per frontier vertex, 1..64 */
    int block;       /* workgroup size: 256, 512 or 1024 */
    int autotuned;   /* 1: per size/degree bucket choices of scc_batch_autotune */
} scc_batch_opts;

typedef struct {
    uint32_t graphs;
    double ms;              /* batch start -> last graph finished */
    double graphs_per_s;
    double p50_ms, p99_ms;  /* per-graph time from batch start to its labels */
    uint64_t vertices, arcs;
    uint32_t launches;      /* kernel launches issued for the batch */
} scc_batch_stats;

/* Copy a batch to the device (stays resident in the context). */
scc_status scc_batch_upload(scc_ctx *ctx, const scc_batch_host *b);
/* All graphs in persistent launches (one per workgroup size in use). */
scc_status scc_batch_run(scc_ctx *ctx, const scc_batch_opts *opts, scc_batch_stats *stats);
/* Baseline: one graph at a time through the single-graph engine with its
 * host-synchronised phase loop (the reference's style of driving). */
scc_status scc_batch_run_serial(scc_ctx *ctx, const scc_opts *opts, scc_batch_stats *stats);
/* Pick lanes, workgroup size and trim/colouring strategy per size and
 * degree bucket by timing each candidate on up to `sample` graphs of the
 * bucket in one launch; a text report of the choices goes to `report`
 * (may be NULL).  Labels of the uploaded batch are overwritten. */
scc_status scc_batch_autotune(scc_ctx *ctx, int sample, char *report, size_t report_len);
/* Labels (row_off layout: V[i]+1 entries at row_off[i]) and SCC counts. */
scc_status scc_batch_labels(scc_ctx *ctx, uint32_t *lab, uint32_t *nscc);

/* ----------------------------------------------------------- serving --- */
/*
 * Continuous-batching request server (no reference counterpart).  Requests
 * (graphs, copied and checked at submit) are admitted while the device
 * bytes of admitted requests stay within `hbm_budget`; each of `slots`
 * worker threads packs up to `max_batch` small requests, in arrival order,
 * into one persistent batched launch on its own stream, so batches overlap
 * on the GPU and requests that arrive meanwhile join the next free slot.  A
 * request with V >= `solo_vertices` (0: 65536) is large: it runs alone on a
 * slot's device-wide engine instead of inside one workgroup, one large
 * request at a time in arrival order, and small requests may overtake it.
 * Latency = submit -> labels on the host.  A request larger than the whole
 * budget is rejected with SCC_ERR_BUDGET.
 */
typedef struct {
    uint64_t submitted, completed, rejected, failed, batches;
    uint64_t solo;                    /* requests run on the device-wide engine */
    double p50_ms, p99_ms, mean_ms;   /* over completed requests */
    double graphs_per_s;              /* completed / (last completion - first arrival) */
    uint64_t peak_bytes;              /* largest admitted footprint */
} scc_server_stats;

scc_status scc_server_start(scc_ctx *ctx, uint64_t hbm_budget, uint32_t max_batch, uint32_t slots,
                            uint32_t solo_vertices, const scc_batch_opts *opts);
scc_status scc_server_submit(scc_ctx *ctx, const scc_host_graph *g, uint64_t *request_id);
/* SCC_ERR_NOTREADY until done; then labels (V+1, may be NULL), SCC count
 * and latency, and the request is released */
scc_status scc_server_poll(scc_ctx *ctx, uint64_t request_id, uint32_t *lab, uint32_t *nscc,
                           double *latency_ms);
scc_status scc_server_wait_all(scc_ctx *ctx);
scc_status scc_server_stats_get(scc_ctx *ctx, scc_server_stats *stats);
scc_status scc_server_stop(scc_ctx *ctx);

/* ------------------------------------------------ device generators --- */
/*
 * The generators above, run on the context's device: the same counter-based
 * draws (so each graph is bit-identical to the host generator's with the
 * same arguments), the CSR and reverse CSR sorted on the device, and the
 * result installed as the context's graph as scc_graph_upload would (no
 * host copy).  scc_gen_communities_dev also installs the community
 * partition and copies it to `part` (V+2 entries) when not NULL.  No
 * reference counterpart: the reference only loads files (load.cpp:12-140).
 */
scc_status scc_gen_er_dev(scc_ctx *ctx, uint32_t V, uint64_t draws, uint32_t cycles, uint32_t cmin,
                          uint32_t cmax, uint64_t seed);
scc_status scc_gen_rmat_dev(scc_ctx *ctx, uint32_t scale, uint64_t E, double a, double b, double c,
                            uint64_t seed);
scc_status scc_gen_grid_dev(scc_ctx *ctx, uint32_t side, double p, uint64_t seed);
scc_status scc_gen_communities_dev(scc_ctx *ctx, uint32_t V, uint32_t K, double deg, double inter,
                                   uint64_t seed, uint32_t *part);
/* The batched mix of scc_gen_batch_mix generated on the device and
 * installed as the context's batch as scc_batch_upload would; with `copy`
 * not NULL the batch is also copied to the host in scc_gen_batch_mix's
 * layout (free with scc_batch_host_free). */
scc_status scc_gen_batch_mix_dev(scc_ctx *ctx, uint32_t num_graphs, uint32_t lgmin, uint32_t lgmax, double dmin,
                                 double dmax, uint64_t seed, scc_batch_host *copy);
/* The context's batch: number of graphs (call with NULL arrays first),
 * then V, E and the row/column offsets of scc_gen_batch_mix's layout. */
scc_status scc_batch_info(scc_ctx *ctx, uint32_t *num_graphs, uint32_t *V, uint32_t *E, uint64_t *row_off,
                          uint64_t *col_off);
/* Size of the context's graph, and a host copy of it (arrays from malloc). */
scc_status scc_graph_size(scc_ctx *ctx, uint32_t *V, uint32_t *E);
scc_status scc_graph_download(scc_ctx *ctx, scc_host_graph *out);

/* Device list layout for V vertices (no reference counterpart; for the
 * layout checks in the tests): out[0..9] = sweep workgroups, trim-1 list
 * slot capacity, trim-1 list words, search workgroups per direction with
 * one and two directions, bottom-up queue slot capacity with one and two
 * directions, queue words, spread slots, narrow-kernel frontier capacity. */
scc_status scc_layout_info(uint32_t V, uint64_t *out, uint32_t n);

/* ------------------------------------------------------- measurement --- */
/* A plain device copy of `bytes` (read + write), `iters` times, on the
 * context's stream; *ms_per_copy = mean duration.  Calibrates the HBM byte
 * counters used to profile the kernels (no reference counterpart). */
scc_status scc_copy_bench(scc_ctx *ctx, uint64_t bytes, int iters, double *ms_per_copy);

/* --------------------------------------------------- drop-in drivers --- */
/*
 * The five drivers of include/scc.h:49-53 with the same parameters and C
 * linkage (for FFI callers: ctypes, cgo, JNI): CSize = E, RSize = V,
 * reference-layout host arrays.  They run on device 0 and print the
 * reference's stdout lines (scc_WCC.cu:948, 966-967; scc_Coloring.cu:669-670;
 * scc_WCC.cu:1576-1577).  They return SCC_OK or an error status and never
 * exit or throw (the reference exits on CUDA errors, scc.h:12-20, and throws
 * on OOM, scc_WCC.cu:618-623).  C++ callers of the reference's own
 * declarations use include/scc_dropin.h, whose void functions with the
 * reference names wrap these.
 */
scc_status scc_vHong(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc,
                     uint32_t *Br, bool t1, bool t2);
scc_status scc_wHong(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc,
                     uint32_t *Br, bool t1, bool t2, int warpSize);
scc_status scc_vSlota(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc,
                      uint32_t *Br, bool t1, bool t2);
scc_status scc_wSlota(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc,
                      uint32_t *Br, bool t1, bool t2, int warpSize);
scc_status scc_detectSCC(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc,
                         uint32_t *Br, uint32_t *Pr, bool t1, bool t2);

#pragma GCC visibility pop
#ifdef __cplusplus
}
#endif
#endif
