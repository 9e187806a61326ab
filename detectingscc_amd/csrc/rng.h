// rng.h -- the counter-based random numbers and seeded permutations of the
// graph generators, as pure functions compiled for both the host (g++) and
// the GPU (hipcc), so a graph generated on the device is bit-identical to
// the one generated on the host.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define SCC_HD __host__ __device__ __forceinline__
#else
#define SCC_HD inline
#endif

namespace sccamd {

// splitmix64 finaliser; rnd(seed, stream, i) is a pure function of its
// arguments, so generated graphs depend neither on the thread count nor on
// where they are generated.
SCC_HD uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

SCC_HD uint64_t rnd(uint64_t seed, uint64_t stream, uint64_t i)
{
    return mix64(mix64(seed + stream * 0x9E3779B97F4A7C15ull) + i * 0xD1B54A32D192ED03ull);
}

// Uniform integer in [0, n) from a 64-bit random value (multiply-shift:
// the high 64 bits of r * n).
SCC_HD uint64_t below(uint64_t r, uint64_t n)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(r, n);
#else
    return (uint64_t)(((unsigned __int128)r * n) >> 64);
#endif
}

// Uniform double in [0, 1) from a 64-bit random value.
SCC_HD double unit(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

// True with probability p for a uniform 64-bit r.
SCC_HD bool bernoulli(uint64_t r, double p) { return p >= 1.0 || unit(r) < p; }

// Seeded permutation of [0, n) as a pure function of the index: a
// four-round balanced Feistel network over the smallest even number of bits
// whose range covers n, with cycle walking (re-encrypting until the value
// falls below n), which restricts a bijection of [0, 2^bits) to one of
// [0, n).  The range is below 4n, so a walk takes fewer than four steps on
// average.  The inverse runs the rounds backwards with the same walk.
struct SeededPerm {
    uint64_t key[4];
    uint32_t n, half;
    uint32_t mask;
};

SCC_HD SeededPerm make_perm(uint32_t n, uint64_t seed)
{
    SeededPerm p;
    uint32_t bits = 2;
    while (bits < 32 && (1ull << bits) < (uint64_t)n) bits += 2;
    p.n = n;
    p.half = bits / 2;
    p.mask = (uint32_t)((1ull << p.half) - 1);
    for (int k = 0; k < 4; k++) p.key[k] = rnd(seed, 9, (uint64_t)k);
    return p;
}

SCC_HD uint32_t feistel_enc(const SeededPerm &p, uint32_t x)
{
    uint32_t L = x >> p.half, R = x & p.mask;
    for (int k = 0; k < 4; k++) {
        const uint32_t f = (uint32_t)mix64(p.key[k] ^ R) & p.mask;
        const uint32_t t = L ^ f;
        L = R;
        R = t;
    }
    return (L << p.half) | R;
}

SCC_HD uint32_t feistel_dec(const SeededPerm &p, uint32_t y)
{
    uint32_t L = y >> p.half, R = y & p.mask;
    for (int k = 3; k >= 0; k--) {
        const uint32_t t = L;
        L = R ^ ((uint32_t)mix64(p.key[k] ^ t) & p.mask);
        R = t;
    }
    return (L << p.half) | R;
}

// i -> perm(i) and its inverse, for i in [0, n)
SCC_HD uint32_t perm_at(const SeededPerm &p, uint32_t i)
{
    uint32_t x = feistel_enc(p, i);
    while (x >= p.n) x = feistel_enc(p, x);
    return x;
}

SCC_HD uint32_t perm_inv(const SeededPerm &p, uint32_t y)
{
    uint32_t x = feistel_dec(p, y);
    while (x >= p.n) x = feistel_dec(p, x);
    return x;
}

// R-MAT quadrant thresholds (16-bit) for probabilities a, b, c.
struct RmatParams {
    uint32_t ta, tab, tabc, scale, words, vcap;
    uint64_t E;
};

// Arc i of an R-MAT draw over [0, 2^scale) with ids >= vcap redrawn (the
// counter advances by E per attempt): writes the unpermuted endpoints.
SCC_HD void rmat_arc(const RmatParams &P, uint64_t seed, uint64_t i, uint32_t &src, uint32_t &dst)
{
    for (uint64_t attempt = 0;; attempt++) {
        const uint64_t ctr = i + attempt * P.E;
        uint32_t r = 0, q = 0, lvl = 0;
        for (uint32_t w = 0; w < P.words; w++) {
            const uint64_t bits = rnd(seed, 10 + w, ctr);
            for (int k = 0; k < 4 && lvl < P.scale; k++, lvl++) {
                const uint32_t x = (uint32_t)(bits >> (16 * k)) & 0xFFFF;
                const uint32_t rb = x >= P.tab, cb = (x >= P.ta && x < P.tab) || x >= P.tabc;
                r = (r << 1) | rb;
                q = (q << 1) | cb;
            }
        }
        if (r < P.vcap && q < P.vcap) {
            src = r;
            dst = q;
            return;
        }
    }
}

} // namespace sccamd
