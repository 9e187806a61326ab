// server.h -- continuous-batching SCC request server on one GPU.
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "scc_amd.h"

namespace sccamd {

class Server {
  public:
    explicit Server(int device) : dev(device) {}
    ~Server();
    scc_status start(uint64_t budget_bytes, uint32_t max_batch, uint32_t slots, uint32_t solo_vertices,
                     const scc_batch_opts &o);
    scc_status submit(const scc_host_graph *g, uint64_t *id);
    scc_status poll(uint64_t id, uint32_t *lab, uint32_t *nscc, double *latency_ms);
    scc_status wait_all();
    scc_status stats(scc_server_stats *s);
    scc_status stop();

    // device bytes a request holds while admitted: in a batch, or alone on
    // the device-wide engine
    static uint64_t footprint(uint32_t V, uint32_t E);
    static uint64_t solo_footprint(uint32_t V, uint32_t E);

  private:
    using clock = std::chrono::steady_clock;
    struct Req {
        uint64_t id = 0;
        uint32_t V = 0, E = 0;
        std::vector<uint32_t> Fr, Fc, Br, Bc, lab;
        uint64_t bytes = 0;
        clock::time_point arrival, done_at;
        uint32_t nscc = 0;
        bool solo = false;   // runs alone on the device-wide engine
        bool done = false;
        scc_status status = SCC_OK;
        std::string error;
    };
    void worker(int slot);
    bool can_admit_locked() const;
    void take_locked(std::vector<Req *> &batch, uint64_t &taken);

    int dev;
    scc_batch_opts opts{};
    scc_opts solo_opts{};
    uint32_t solo_vertices = 1u << 16;
    uint64_t budget = 0, inflight = 0, peak = 0;
    uint32_t solo_running = 0;         // slots running a large request
    uint32_t solo_limit = 1;           // at most slots - 1 of them (at least 1)
    uint32_t max_batch = 64;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    std::vector<std::thread> workers;
    bool running = false, stopping = false;
    uint64_t next_id = 1, submitted = 0, completed = 0, rejected = 0, batches = 0, failed = 0, solo_runs = 0;
    std::vector<double> latencies;
    clock::time_point first_arrival, last_done;
    double busy_ms = 0;
};

} // namespace sccamd
