// batch_engine.h -- host side of the batched persistent form.
#pragma once
#include <string>
#include <vector>

#include "batch.h"
#include "scc_amd.h"

namespace sccamd {

class Engine;

class BatchEngine {
  public:
    ~BatchEngine();
    scc_status upload(int device, hipStream_t stream, const scc_batch_host *b);
    // a batch in device memory in the host layout (row_off/col_off given on
    // the host, d_roff = row_off on the device), valid by construction
    scc_status upload_device(int device, hipStream_t stream, uint32_t ng, const uint32_t *Vs, const uint32_t *Es,
                             const uint64_t *roff, const uint64_t *coff, const uint64_t *d_roff,
                             const uint32_t *dFr, const uint32_t *dBr, const uint32_t *dFc, const uint32_t *dBc);
    scc_status run(const scc_batch_opts &o, scc_batch_stats *out);
    scc_status run_serial(Engine &eng, const scc_opts &o, scc_batch_stats *out);
    scc_status autotune(int sample, std::string *report);
    scc_status labels(uint32_t *lab, uint32_t *nscc);
    // the uploaded batch's sizes and host-layout offsets (arrays may be NULL)
    uint32_t info(uint32_t *Vs, uint32_t *Es, uint64_t *roff, uint64_t *coff) const;

    static constexpr int NBUCKETS = 12;

  private:
    void release();
    scc_status prepare(int device, hipStream_t stream, uint32_t ng, const uint32_t *Vs, const uint32_t *Es,
                       const uint64_t *roff, const uint64_t *coff);
    scc_status finish_upload();
    scc_status launch(const std::vector<uint32_t> &idx, const std::vector<uint32_t> &cfg,
                      const std::vector<int> &block, double *ms, std::vector<double> *lat_ms,
                      uint32_t *launches);

    int dev = 0;
    hipStream_t s1 = nullptr, stream2 = nullptr;
    BatchDev d;
    uint32_t G = 0;
    std::vector<uint32_t> V, E;
    std::vector<uint64_t> row_off, voff, eoff;
    std::vector<int> bucket;
    std::vector<uint32_t> best_cfg;
    std::vector<int> best_block;
    std::vector<void *> allocs;
    uint64_t capS = 0, capA = 0;
    uint32_t capG = 0;
    uint64_t *m_voff = nullptr, *m_eoff = nullptr;
    uint32_t *m_V = nullptr, *m_Fr = nullptr, *m_Br = nullptr, *m_Fc = nullptr, *m_Bc = nullptr;
    uint32_t *m_cfg = nullptr, *m_order = nullptr;
    std::vector<uint32_t> stage;        // host staging of the row arrays
    uint64_t nverts = 0, narcs = 0;
    double ticks_per_ms = 100000.0;
    int num_cus = 256;
    bool tuned = false, ran = false;
};

} // namespace sccamd
