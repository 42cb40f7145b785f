/*
 * Counter-based SplitMix64 random stream shared by the workload generators,
 * the Python mirror (hpc_amd/workloads.py) and the tests.
 *
 * draw(seed, stream, i) is a pure function of its three arguments, so a value
 * can be produced by any thread in any order and the C, HIP and numpy
 * generators agree bit for bit.  The mixing function is Steele, Lea and
 * Flood's SplitMix64 finaliser (OOPSLA 2014) with the constants published
 * there; uniform doubles take the top 53 bits.
 */
#ifndef XJPCG_RNG_H
#define XJPCG_RNG_H

#include <stdint.h>

#ifdef __HIPCC__
#define XJPCG_HD __host__ __device__ inline
#else
#define XJPCG_HD static inline
#endif

#define XJPCG_GOLDEN 0x9E3779B97F4A7C15ULL

XJPCG_HD uint64_t xjpcg_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* i-th 64-bit draw of stream `stream` under `seed`. */
XJPCG_HD uint64_t xjpcg_draw(uint64_t seed, uint64_t stream, uint64_t i) {
    uint64_t key = xjpcg_mix64(seed + XJPCG_GOLDEN * (stream + 1));
    return xjpcg_mix64(key + XJPCG_GOLDEN * (i + 1));
}

/* uniform in [0, 1) with 53 random bits */
XJPCG_HD double xjpcg_u01(uint64_t r) { return (double)(r >> 11) * 0x1.0p-53; }

/* uniform in [lo, hi) */
XJPCG_HD double xjpcg_uniform(uint64_t seed, uint64_t stream, uint64_t i, double lo, double hi) {
    return lo + (hi - lo) * xjpcg_u01(xjpcg_draw(seed, stream, i));
}

#endif
