// Deterministic double-precision log, log1p and exp for the seeded
// generators: the same bits on the host and on the device, so a corpus
// generated on the GPU is the corpus the host generator writes. Every
// operation is a single IEEE add, multiply or divide (correctly rounded on
// both sides) with contraction into fused multiply-adds switched off; no libm
// call whose last-bit behaviour differs between libraries.
//
// log: x = m 2^e with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
// s = (m - 1) / (m + 1), |s| <= 0.1716, summed as the odd series
// 2 s (1 + s^2/3 + s^4/5 + ... + s^22/23) (truncation below 1e-17 relative).
// exp: x = k ln2 + r, |r| <= ln2 / 2, Taylor series of e^r to r^16 / 16!,
// scaled by 2^k through the exponent field.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace stormhip {

__host__ __device__ inline double det_bits_to_double(uint64_t b) {
    union { uint64_t u; double d; } v;
    v.u = b;
    return v.d;
}
__host__ __device__ inline uint64_t det_double_to_bits(double d) {
    union { uint64_t u; double d; } v;
    v.d = d;
    return v.u;
}

// 2 atanh(s) for |s| <= 0.18.
__host__ __device__ inline double det_atanh2(double s) {
#pragma clang fp contract(off)
    const double z = s * s;
    double p = 1.0 / 23.0;
    p = p * z + 1.0 / 21.0;
    p = p * z + 1.0 / 19.0;
    p = p * z + 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    const double s2 = s + s;
    return s2 + s2 * (z * p);
}

// Natural log of a positive normal double.
__host__ __device__ inline double det_log(double x) {
#pragma clang fp contract(off)
    const uint64_t b = det_double_to_bits(x);
    int e = (int)((b >> 52) & 0x7FF) - 1023;
    double m = det_bits_to_double((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 1.4142135623730951) {   // m in (sqrt 2, 2): halve it
        m = m * 0.5;
        e += 1;
    }
    const double f = m - 1.0;
    const double s = f / (m + 1.0);
    const double ln2_hi = 6.93147180369123816490e-01;   // high bits of ln 2 (e * hi is exact)
    const double ln2_lo = 1.90821492927058770002e-10;
    const double de = (double)e;
    return de * ln2_hi + (det_atanh2(s) + de * ln2_lo);
}

// log(1 + y) for y > -1, accurate for small |y|.
__host__ __device__ inline double det_log1p(double y) {
#pragma clang fp contract(off)
    if (y > -0.25 && y < 0.25) return det_atanh2(y / (2.0 + y));
    return det_log(1.0 + y);
}

// e^x for x in [-700, 700].
__host__ __device__ inline double det_exp(double x) {
#pragma clang fp contract(off)
    const double inv_ln2 = 1.4426950408889634;
    const double t = x * inv_ln2;
    const double kd = t >= 0.0 ? (double)(int64_t)(t + 0.5) : -(double)(int64_t)(0.5 - t);
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double r = (x - kd * ln2_hi) - kd * ln2_lo;
    double p = 1.0 / 20922789888000.0;          // 1/16!
    p = p * r + 1.0 / 1307674368000.0;          // 1/15!
    p = p * r + 1.0 / 87178291200.0;
    p = p * r + 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    const int k = (int)kd;
    return p * det_bits_to_double((uint64_t)(k + 1023) << 52);
}

// splitmix64 stream (kernels/storm_gen.h:84-102): next, below, unit.
struct DetRng {
    uint64_t s;
    __host__ __device__ explicit DetRng(uint64_t seed) : s(seed ? seed : 0x9E3779B97F4A7C15ull) {}
    __host__ __device__ uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    __host__ __device__ uint32_t below(uint32_t n) {
        if (n == 0) return 0;
        const uint64_t lim = ~0ull - (~0ull % n);
        uint64_t r;
        do { r = next(); } while (r >= lim);
        return (uint32_t)(r % n);
    }
    __host__ __device__ double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Seed of row r's own stream: draw r of Rng(seed).
__host__ __device__ inline uint64_t det_row_seed(uint64_t seed, uint64_t r) {
    uint64_t z = (seed ? seed : 0x9E3779B97F4A7C15ull) + (r + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Geometric on {1, 2, ...} with the given mean (1 when mean <= 1).
__host__ __device__ inline uint32_t det_geometric(DetRng& g, double mean) {
#pragma clang fp contract(off)
    if (mean <= 1.0) return 1;
    double u = g.unit();
    if (u <= 0.0) u = 1e-300;
    const double x = __builtin_floor(det_log(u) / det_log1p(-1.0 / mean)) + 1.0;
    return x < 4.0e9 ? (uint32_t)x : 4000000000u;
}

// Poisson(lambda): inversion by sequential search on one uniform per draw of
// mean at most 30; a larger mean is the sum of such draws (Poisson(a + b) =
// Poisson(a) + Poisson(b)), so e^-mean never underflows.
__host__ __device__ inline uint32_t det_poisson(DetRng& g, double lambda) {
#pragma clang fp contract(off)
    uint32_t total = 0;
    while (lambda > 0.0) {
        const double l = lambda > 30.0 ? 30.0 : lambda;
        lambda = lambda - l;
        const double u = g.unit();
        double p = det_exp(-l), c = p;
        uint32_t k = 0;
        while (u >= c && k < 1000) {
            ++k;
            p = p * (l / (double)k);
            c = c + p;
        }
        total += k;
    }
    return total;
}

// Bernoulli(d) row by geometric skips: the next set position after `pos`
// (pos = -1 before the first); returns a value >= universe at the end.
__host__ __device__ inline double det_bernoulli_next(DetRng& g, double l1, double pos) {
#pragma clang fp contract(off)
    double v = g.unit();
    if (v <= 0.0) v = 1e-300;
    return pos + (__builtin_floor(det_log(v) / l1) + 1.0);
}

// One row of the run-heavy generator: Poisson(lambda) runs at uniform starts
// with geometric lengths of mean `mean_run` (clipped at the universe), then
// round(frac * run bits) singletons at uniform positions (with repetition;
// the row is the union). emit(start, end) receives each half-open interval in
// draw order.
template <class Emit>
__host__ __device__ inline void det_runs_row(DetRng& g, uint32_t u, double lambda,
                                             double mean_run, double frac, Emit&& emit) {
#pragma clang fp contract(off)
    const uint32_t nruns = det_poisson(g, lambda);
    uint64_t bits = 0;
    for (uint32_t q = 0; q < nruns; ++q) {
        const uint32_t s = g.below(u);
        const uint32_t g_len = det_geometric(g, mean_run);
        const uint32_t len = g_len < u - s ? g_len : u - s;
        emit(s, s + len);
        bits += len;
    }
    const uint64_t extra = (uint64_t)__builtin_floor(frac * (double)bits + 0.5);
    for (uint64_t q = 0; q < extra; ++q) {
        const uint32_t x = g.below(u);
        emit(x, x + 1);
    }
}

}  // namespace stormhip
