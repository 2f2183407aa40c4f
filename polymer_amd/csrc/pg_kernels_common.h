// pg_kernels_common.h -- deterministic synthetic-input functions shared by the
// device generators.  They are functions of (seed, vertex pair) only, so a
// graph's weights do not depend on edge order; the float ones are compiled
// with contraction off so every operation rounds as written (no fma), as the
// oracle's C code does.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace pg {

// splitmix64 finaliser (Steele, Lea, Flood 2014)
__host__ __device__ __forceinline__ uint64_t dmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t host_mix64(uint64_t z) { return dmix64(z); }

inline void host_rmat_thresholds(double a, double b, double c, uint32_t thr[3]) {
    const double t0 = std::floor(a * 4294967296.0);
    const double t1 = std::floor((a + b) * 4294967296.0);
    const double t2 = std::floor((a + b + c) * 4294967296.0);
    thr[0] = t0 >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)t0;
    thr[1] = t1 >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)t1;
    thr[2] = t2 >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)t2;
}

// int32 weight in [1,255], symmetric in (u,v)
__device__ __forceinline__ int32_t dweight_i32(uint64_t seed, int32_t u, int32_t v) {
    const uint32_t lo = (uint32_t)(u < v ? u : v), hi = (uint32_t)(u < v ? v : u);
    const uint64_t h = dmix64(dmix64(seed ^ 0xC2B2AE3D27D4EB4Full) ^ (((uint64_t)lo << 32) | hi));
    return (int32_t)(1 + (h >> 33) % 255u);
}

// fp32 value (k+1) 2^-24 in (0,1], directed pair
__device__ __forceinline__ float dvalue_f32(uint64_t seed, int32_t u, int32_t v) {
    const uint64_t h = dmix64(dmix64(seed ^ 0x165667B19E3779F9ull) ^
                              (((uint64_t)(uint32_t)u << 32) | (uint32_t)v));
    return (float)((h >> 40) + 1) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ float dunit_f32(uint64_t seed, uint64_t i) {
    const uint64_t h = dmix64(dmix64(seed ^ 0x27D4EB2F165667C5ull) + i);
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ float drange_f32(uint64_t seed, uint64_t i, float lo, float hi) {
    // __fmul_rn/__fadd_rn are plain * and + here: forbid their fusion into an fma
#pragma clang fp contract(off)
    const float u = dunit_f32(seed, i);
    const float span = hi - lo;
    const float t = span * u;
    return lo + t;
}

// BP edge potential W_e[k/2][k%2] for the edge s->d
__device__ __forceinline__ float dbp_pot(uint64_t seed, int32_t s, int32_t d, int k, float lo,
                                         float hi) {
    const uint64_t i = dmix64(((uint64_t)(uint32_t)s << 32) | (uint32_t)d) + (uint64_t)k;
    return drange_f32(seed ^ 0xB5297A4D3F84D5B5ull, i, lo, hi);
}

}  // namespace pg
