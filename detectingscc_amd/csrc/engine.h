// engine.h -- the device-resident SCC engine behind the C ABI.
#pragma once
#include <memory>
#include <vector>

#include "scc_amd.h"
#include "scc_dev.h"

struct scc_ctx;

namespace sccamd {

const char *phase_name(int p);
// Looks up this device's address of the allocation-free entry point's
// scratch array (noalloc.hip); called once per device at context creation.
scc_status noalloc_init(int device);

class Engine {
  public:
    ~Engine();
    scc_status init(int device);
    // `checked`: the caller has already validated the arrays (the server's
    // submit does), so the host pass over them is skipped
    scc_status upload(uint32_t V, uint32_t E, const uint32_t *Fr, const uint32_t *Fc,
                      const uint32_t *Br, const uint32_t *Bc, bool checked = false);
    scc_status upload_partition(const uint32_t *part);
    scc_status upload_device(uint32_t V, uint32_t E, const uint32_t *dFr, const uint32_t *dFc,
                             const uint32_t *dBr, const uint32_t *dBc, hipStream_t from);
    // A graph built in place on the device (gen_dev.hip): adopt_begin sizes
    // the arrays for V vertices and E arcs, the caller fills g.Fr/Fc/Br/Bc in
    // the reference layout, adopt_end makes it the uploaded graph.
    scc_status adopt_begin(uint32_t V, uint32_t E);
    void adopt_end();
    scc_status run(const scc_opts &o, scc_stats *out);
    scc_status phase_local(uint32_t *out);
    scc_status labels(uint32_t *lab);
    // size histogram of the last run's labels, computed on the device:
    // (size, count) pairs ascending by size
    scc_status histogram(uint64_t *pairs, uint32_t cap, uint32_t *npairs, double *ms);
    scc_status condense(scc_host_graph *out, uint32_t *node_of);

    scc_status phase_trim1(const uint8_t *dead, uint8_t *removed);
    scc_status phase_trim2(const uint8_t *dead, uint32_t *pair);
    scc_status phase_first_pivot(const uint8_t *dead, uint32_t *pivot);
    scc_status phase_reach(uint32_t src, int dir, int expand, int W, const uint8_t *dead,
                           uint8_t *out, uint32_t *levels, double *ms);
    scc_status phase_min_label(const uint8_t *dead, uint32_t *out);
    scc_status phase_max_color(const uint8_t *dead, uint32_t *out);

    int device = -1;
    hipStream_t stream = nullptr;
    DevGraph g;
    DevState st;
    bool has_graph = false, has_part = false, ran = false;

  private:
    void release();
    scc_status reserve(uint32_t V, uint32_t E);
    scc_status read_ctr();                 // publish_ctr + wait_ctr
    uint32_t publish_ctr();                // queue a publication of the counters, returns its number
    scc_status wait_ctr(uint32_t seq);     // wait for publication `seq`, copy it into h_ctr
    void zero_ctr(int first, int n);
    // A device timestamp opening `phase` (-1: untimed, also the start and
    // end of a decomposition).  Phase marks are recorded only with
    // phase_marks (each costs the GPU a few microseconds) or `force`.
    void mark(int phase, bool force = false);
    void collect_times(scc_stats *s);
    uint32_t bpd(int W) const;
    scc_status reach(bool fw, bool bw, bool single, int expand, int W, uint32_t *levels_out);
    scc_status reach_do(bool fw, bool bw, bool single, uint32_t *levels_out);
    DirArgs dir_args(int d, uint32_t level) const;
    // single: every live vertex has class 0 (the first trim of a
    // decomposition); none_dead: besides, the state is fresh (no vertex dead,
    // all in class 0, labels v | LAB_FINAL; its per-vertex words may be unwritten;
    // every counter zero: the state initialisation was the last thing queued)
    // wait = false: one queued batch of passes, no look at the counters
    scc_status trim1(bool fixpoint, bool single = false, bool none_dead = false, bool wait = true,
                     bool labels = true);   // labels = false: a fresh state's labels stay unwritten
    void init_vis(bool elect_first = false, bool use_live_bits = false);
    // the single-class trim-1's live bitmap (in d_tmp8)
    uint32_t *live_bits() const { return reinterpret_cast<uint32_t *>(d_tmp8); }
    uint64_t limit() const;
    scc_status no_convergence(const char *what);
    // canon >= 0: queue the canonical labels (with sizes if canon == 1) gated
    // on no live vertex being left, before the look at the counters
    // use_live_bits: the first round reads live_bits() (exact, every live vertex
    // in class 0: a single-class trim-1 just ran) instead of the class words
    // lazy: the first round after a label-less trim (k_update lazy, labels
    // from the bitmaps)
    scc_status pivot_round(bool first, int expand, int W, scc_stats *stats, uint32_t *live, int canon = -1,
                           bool use_live_bits = false, bool lazy = false);
    scc_status min_label();
    scc_status color_loop(int expand, int W, scc_stats *stats);
    scc_status detect_phases(const scc_opts &o, int expand, int W, scc_stats *stats, bool local_only);
    struct DevBuf {
        void *p = nullptr;
        size_t cap = 0;
    };
    template <class T> scc_status scratch(DevBuf &b, size_t n, T **out);
    scc_status phase_begin(const uint8_t *dead);
    scc_status fetch8(uint8_t *out);
    scc_status fetch32(uint32_t *out, const uint32_t *src);

    uint32_t capV = 0, capE = 0;
    size_t qf_words = 0, qb_words = 0;   // allocated words of each FW / BW queue
    uint32_t *h_ctr = nullptr;        // pinned mirror of st.ctr
    uint32_t *h_pub = nullptr;        // mapped coherent host memory the GPU publishes counters into (2 buffers)
    uint32_t *d_pub = nullptr;        // its device address
    uint32_t pub_seq = 0;             // last publication requested
    uint8_t *d_tmp8 = nullptr;        // V+1 bytes for phase inputs/outputs; during a decomposition the live bitmap
    int check_every = 0;
    bool search_stats = false;        // count |R| and |E_R| of the first search
    bool do_seeded = false;           // init_vis(true) wrote the level-0 state of a two-direction search
    bool phase_marks = true;          // scc_opts.phase_times of the current run (phase APIs: on)
    uint32_t trim_epoch = 0;          // last epoch stamped by a trim-1 list pass
    bool trace_levels = false;        // print each reachability decision to stderr
    std::vector<hipEvent_t> events;
    std::vector<int> mark_phase;
    size_t nmarks = 0;
    int first_reach_marks[2] = {-1, -1};
    uint32_t num_parts = 0;
    DevBuf buf[16];                    // detectSCC scratch (lazily sized)
    std::unique_ptr<Engine> sub;       // engine for the quotient graph
};

} // namespace sccamd

namespace sccamd {
class BatchEngine;
class Server;
}

// A context: the single-graph engine, plus the batched form and its
// one-graph-per-launch baseline engine when used.  Destroyed in capi.hip,
// which sees the complete types.
struct scc_ctx {
    sccamd::Engine eng;
    std::unique_ptr<sccamd::BatchEngine> batch;
    std::unique_ptr<sccamd::Engine> serial;
    std::unique_ptr<sccamd::Server> server;
    ~scc_ctx();
};
