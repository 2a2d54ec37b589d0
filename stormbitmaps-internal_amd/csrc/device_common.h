// Device helpers shared by the engine's kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace stormhip {

// Packed upper-triangle index of pair (i, j), i < j < n, original row order:
// row i starts after the sum of the lengths of rows 0..i-1.
__host__ __device__ inline uint64_t tri_index(uint64_t i, uint64_t j, uint64_t n) {
    return i * n - i * (i + 1) / 2 + (j - i - 1);
}

// Linear id -> (bi, bj), bj >= bi, row-major over the upper triangle of a
// T x T tile grid. Row bi starts at bi*T - bi*(bi-1)/2.
__host__ __device__ inline void tri_tile(uint32_t t, uint32_t T, uint32_t& bi, uint32_t& bj) {
    // Solve start(bi) <= t with start(b) = b*(2T - b + 1)/2, then correct.
    const double tt = (double)T + 0.5;
    double disc = tt * tt - 2.0 * (double)t;
    if (disc < 0) disc = 0;
    int64_t b = (int64_t)(tt - sqrt(disc));
    if (b < 0) b = 0;
    if (b > (int64_t)T - 1) b = T - 1;
    auto start = [T](int64_t x) { return (uint64_t)(x * (2 * (int64_t)T - x + 1) / 2); };
    while (b > 0 && start(b) > t) --b;
    while (b + 1 < (int64_t)T && start(b + 1) <= t) ++b;
    bi = (uint32_t)b;
    bj = (uint32_t)(b + (t - start(b)));
}

// (bi, bj), bi <= bj < T -> the row-major id that tri_tile inverts.
__host__ __device__ inline uint64_t tri_tile_id(uint32_t bi, uint32_t bj, uint32_t T) {
    return (uint64_t)bi * (2ull * T - bi + 1) / 2 + (bj - bi);
}

// Grid block -> position in a tile order. Workgroups are dispatched to the
// 8 XCDs round-robin (block b on XCD b % 8), so block b takes position
// (b % 8) * (G8 / 8) + b / 8 of the first G8 = 8 * floor(G / 8) positions:
// each XCD walks a contiguous range of the order, and tiles running together
// on one XCD are neighbours in it. The tail keeps its own order. A bijection
// on [0, G).
__host__ __device__ inline uint32_t xcd_position(uint32_t b, uint32_t G) {
    const uint32_t G8 = G / 8 * 8;
    if (b >= G8) return b;
    return (b % 8) * (G8 / 8) + b / 8;
}

// Banded order of the upper-triangle tiles of a T x T grid: bands of `gh`
// tile rows; inside a band, column by column (bj ascending), the band's rows
// bi <= bj. Neighbours in this order share row panels both ways (a run of 32
// tiles touches about gh + 32/gh panels instead of 33).
__host__ __device__ inline void banded_tile(uint32_t v, uint32_t T, uint32_t gh, uint32_t& bi,
                                            uint32_t& bj) {
    uint64_t base = 0;
    uint32_t lo = 0, h = 0;
    for (lo = 0; lo < T; lo += gh) {
        h = gh < T - lo ? gh : T - lo;
        const uint64_t cnt = (uint64_t)h * (h + 1) / 2 + (uint64_t)h * (T - lo - h);
        if (v < base + cnt) break;
        base += cnt;
    }
    uint64_t off = v - base;
    const uint64_t dsz = (uint64_t)h * (h + 1) / 2;
    if (off < dsz) {
        // Diagonal block: column c holds rows lo .. lo + c.
        uint32_t c = 0;
        while ((uint64_t)(c + 1) * (c + 2) / 2 <= off) ++c;
        bj = lo + c;
        bi = lo + (uint32_t)(off - (uint64_t)c * (c + 1) / 2);
    } else {
        off -= dsz;
        bj = lo + h + (uint32_t)(off / h);
        bi = lo + (uint32_t)(off % h);
    }
}

__device__ inline uint64_t wave_sum_u64(uint64_t v) {
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, off, 64);
        const uint32_t hi = __shfl_xor((uint32_t)(v >> 32), off, 64);
        v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}

__device__ inline uint32_t wave_sum_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// splitmix64 finaliser (the output function of kernels/storm_gen.h:84-102).
__host__ __device__ inline uint64_t splitmix_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

}  // namespace stormhip
