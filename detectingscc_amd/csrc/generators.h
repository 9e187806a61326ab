// generators.h -- arc-list generators shared by the C-ABI generators, the
// batched mix and the serving stream.  Arcs are appended to (s, d).
#pragma once
#include <cstdint>
#include <vector>

namespace sccamd {

std::vector<uint32_t> seeded_permutation(uint32_t n, uint64_t seed);
void er_arcs(uint32_t V, uint64_t draws, uint64_t seed, std::vector<uint32_t> &s,
             std::vector<uint32_t> &d);
void planted_cycles(uint32_t V, uint32_t cycles, uint32_t cmin, uint32_t cmax, uint64_t seed,
                    std::vector<uint32_t> &s, std::vector<uint32_t> &d);
void rmat_arcs(uint32_t scale, uint32_t vcap, uint64_t E, double a, double b, double c,
               uint64_t seed, const std::vector<uint32_t> *perm, std::vector<uint32_t> &s,
               std::vector<uint32_t> &d);
void grid_arcs(uint32_t side, double p, uint64_t seed, std::vector<uint32_t> &s,
               std::vector<uint32_t> &d);

} // namespace sccamd
