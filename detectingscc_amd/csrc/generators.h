// generators.h -- arc-list generators shared by the C-ABI generators, the
// batched mix and the serving stream.  Arcs are appended to (s, d).
#pragma once
#include <cstdint>
#include <vector>

#include "rng.h"

namespace sccamd {

void er_arcs(uint32_t V, uint64_t draws, uint64_t seed, std::vector<uint32_t> &s,
             std::vector<uint32_t> &d, int threads = 0);
void planted_cycles(uint32_t V, uint32_t cycles, uint32_t cmin, uint32_t cmax, uint64_t seed,
                    std::vector<uint32_t> &s, std::vector<uint32_t> &d);
RmatParams rmat_params(uint32_t scale, uint32_t vcap, uint64_t E, double a, double b, double c);
void rmat_arcs(uint32_t scale, uint32_t vcap, uint64_t E, double a, double b, double c,
               uint64_t seed, bool permute, std::vector<uint32_t> &s, std::vector<uint32_t> &d,
               int threads = 0);
void grid_arcs(uint32_t side, double p, uint64_t seed, std::vector<uint32_t> &s,
               std::vector<uint32_t> &d);
void csr_build_serial(uint32_t V, uint64_t E, const uint32_t *src, const uint32_t *dst, uint32_t *Fr,
                      uint32_t *Fc, uint32_t *Br, uint32_t *Bc);

// The per-graph draws of the batched mix that fix a graph's size and kind
// (graph seed gs, V, target arc count E, kind), shared by the host and
// device generators so both build the same graphs.
struct MixParams {
    uint64_t gs;
    uint32_t V;
    uint64_t E;
    uint8_t kind;
};
MixParams mix_params(uint64_t seed, uint32_t index, uint32_t lgmin, uint32_t lgmax, double dmin, double dmax);
// planted-cycle graphs: cycle count and each cycle's length
uint32_t mix_cycles(const MixParams &m);
uint32_t mix_cycle_len(const MixParams &m, uint32_t j);
constexpr uint64_t kMixFillSalt = 0x5bd1e995ull;   // seed of the ER fill of a planted-cycle graph: gs ^ salt

// One graph of the batched mix: kind 0 ER, 1 R-MAT, 2 planted cycles + ER.
struct MixGraph {
    uint32_t V;
    uint8_t kind;
    std::vector<uint32_t> s, d;
};
MixGraph mix_graph(uint64_t seed, uint32_t index, uint32_t lgmin, uint32_t lgmax, double dmin, double dmax);
// One request of the serving stream: V lognormal (median, sigma) clamped
// to [2, vcap], mean degree `degree`, kind 0 ER or 1 R-MAT (even odds).
MixGraph serving_graph(uint64_t seed, uint32_t index, double median_v, double sigma, uint32_t vcap,
                       double degree);

} // namespace sccamd
