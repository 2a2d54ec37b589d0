// Device-side corpus generation, representations built from bitmaps or runs,
// the rank9 / zone-map indexes, and the column closed forms.
//
//   generation   the seeded workloads written straight into device memory in
//                the representation they are stored in: Bernoulli rows as
//                bitmaps (one draw per bit of the stream Rng(seed), as the
//                dense corpora) or as lists (geometric skips on the row's own
//                stream), run-heavy rows as maximal runs. The row streams and
//                transforms are those of the host generator (gen_host.hip,
//                detmath.h), so a row generated here equals the host's.
//   extraction   a scratch bitmap per row -> maximal runs or a position list
//                (two passes: counts, then writes at scanned offsets).
//   meta         RowMeta (kernels/storm_repr.h:213-231) of rows held only as
//                runs or only as bitmaps, with build_row's index gates
//                (kernels/storm_repr.cpp:184-190).
//   rank / occ   build_rank (kernels/storm_repr.cpp:25-45) and build_occ
//                (:47-59) for the bitmap rows whose gates pass.
//   closed form  column counts c_b of a corpus, then sum_b C(c_b, 2) (the sum
//                over pairs of |Xi n Xj|) and every row's sum_{b in Xi}
//                (c_b - 1): an exact check that shares nothing with the cells.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "build.h"
#include "detmath.h"
#include "device_common.h"
#include "general.h"
#include "internal.h"

namespace stormhip {

namespace {

// ---- block helpers (256 threads) -------------------------------------------

// Exclusive prefix of v over the block; *total receives the block sum.
__device__ uint64_t block_scan_u64(uint64_t v, uint64_t* total) {
    __shared__ uint64_t part[4];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t lo = __shfl_up((uint32_t)x, d, 64);
        const uint32_t hi = __shfl_up((uint32_t)(x >> 32), d, 64);
        if (lane >= (uint32_t)d) x += ((uint64_t)hi << 32) | lo;
    }
    __syncthreads();
    if (lane == 63) part[wave] = x;
    __syncthreads();
    uint64_t before = 0, all = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
        if (w < wave) before += part[w];
        all += part[w];
    }
    *total = all;
    return before + x - v;
}

__device__ uint64_t block_sum_u64(uint64_t v) {
    uint64_t t;
    (void)block_scan_u64(v, &t);
    return t;
}

// ---- generation --------------------------------------------------------------

// Bernoulli(p) bits of rows [row0, row0 + rows) into bitmap slots
// slot0 + (r - row0): bit b of row r is draw r * universe + b of Rng(seed)
// (the dense corpora's stream), set when (next() >> 11) < ceil(p * 2^53).
__global__ void gen_bits_kernel(uint64_t* __restrict__ store, uint32_t ld, uint32_t slot0,
                                uint32_t row0, uint32_t rows, uint32_t universe, uint64_t thr,
                                uint64_t seed) {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (uint64_t)rows * ld) return;
    const uint32_t k = (uint32_t)(gid / ld), w = (uint32_t)(gid % ld);
    const uint64_t r = row0 + k;
    uint64_t word = 0;
    if ((uint64_t)w * 64 < universe) {
        const uint32_t nb = min(64u, universe - w * 64);
        uint64_t s = seed + (r * universe + (uint64_t)w * 64) * kGolden;
        for (uint32_t b = 0; b < nb; ++b) {
            s += kGolden;
            if ((splitmix_mix(s) >> 11) < thr) word |= 1ull << b;
        }
    }
    store[(uint64_t)(slot0 + k) * ld + w] = word;
}

// One thread per row: its own stream det_row_seed(seed, r) drives the row's
// draws; set positions go into the row's (zeroed) scratch bitmap.
__global__ void gen_skip_kernel(uint64_t* __restrict__ scratch, uint32_t ld, uint32_t row0,
                                uint32_t rows, uint32_t universe, double d, uint64_t seed) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= rows) return;
    uint64_t* w = scratch + (uint64_t)k * ld;
    DetRng g(det_row_seed(seed, row0 + k));
    if (d <= 0.0) return;
    if (d >= 1.0) {
        for (uint32_t p = 0; p < universe; ++p) w[p >> 6] |= 1ull << (p & 63);
        return;
    }
    const double l1 = det_log1p(-d);
    for (double pos = det_bernoulli_next(g, l1, -1.0); pos < (double)universe;
         pos = det_bernoulli_next(g, l1, pos)) {
        const uint32_t p = (uint32_t)pos;
        w[p >> 6] |= 1ull << (p & 63);
    }
}

__device__ inline void set_range(uint64_t* w, uint32_t s, uint32_t e) {
    if (s >= e) return;
    const uint32_t ws = s >> 6, we = (e - 1) >> 6;
    const uint64_t lo = ~0ull << (s & 63), hi = ~0ull >> (63 - ((e - 1) & 63));
    if (ws == we) { w[ws] |= lo & hi; return; }
    w[ws] |= lo;
    for (uint32_t q = ws + 1; q < we; ++q) w[q] = ~0ull;
    w[we] |= hi;
}

__global__ void gen_runs_kernel(uint64_t* __restrict__ scratch, uint32_t ld, uint32_t row0,
                                uint32_t rows, uint32_t universe, double lambda, double mean_run,
                                double frac, uint64_t seed) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= rows) return;
    uint64_t* w = scratch + (uint64_t)k * ld;
    DetRng g(det_row_seed(seed, row0 + k));
    det_runs_row(g, universe, lambda, mean_run, frac,
                 [&](uint32_t s, uint32_t e) { set_range(w, s, e); });
}

__global__ void zero_u64_kernel(uint64_t* p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// ---- extraction: bitmap rows -> runs / lists -----------------------------------
//
// A workgroup per row; thread t owns words [t*span, (t+1)*span). A run starts
// at p when bit p is set and bit p-1 is not, and ends at p (exclusive) when
// bit p-1 is set and bit p is not; the row's k-th start pairs with its k-th
// end. Bits past the universe are zero, and an end at the universe itself is
// added by the last word's owner. mode 0: runs, 1: positions.
__device__ inline void word_edges(const uint64_t* w, uint32_t q, uint64_t& starts, uint64_t& ends) {
    const uint64_t x = w[q];
    const uint64_t prev = q ? (w[q - 1] >> 63) : 0;
    const uint64_t shifted = (x << 1) | prev;    // bit p = bit p-1 of the row
    starts = x & ~shifted;
    ends = ~x & shifted;
}

__global__ __launch_bounds__(256) void extract_kernel(const uint64_t* __restrict__ scratch,
                                                      uint32_t ld, uint32_t nw, uint32_t universe,
                                                      uint32_t rows, int mode, int pass,
                                                      uint64_t* __restrict__ counts,
                                                      const uint64_t* __restrict__ offs,
                                                      uint32_t* __restrict__ out_a,
                                                      uint32_t* __restrict__ out_b) {
    const uint32_t k = blockIdx.x;
    if (k >= rows) return;
    const uint64_t* w = scratch + (uint64_t)k * ld;
    const uint32_t span = (nw + blockDim.x - 1) / blockDim.x;
    const uint32_t q0 = min(nw, threadIdx.x * span), q1 = min(nw, q0 + span);
    const bool tail_end = nw && (universe == nw * 64u) && threadIdx.x == (nw - 1) / span;
    uint64_t mine = 0, mine_end = 0;
    for (uint32_t q = q0; q < q1; ++q) {
        if (mode == 0) {
            uint64_t s, e;
            word_edges(w, q, s, e);
            mine += __popcll(s);
            mine_end += __popcll(e);
        } else {
            mine += __popcll(w[q]);
        }
    }
    if (mode == 0 && tail_end && (w[nw - 1] >> 63)) mine_end += 1;
    uint64_t total;
    const uint64_t before = block_scan_u64(mine, &total);
    if (pass == 0) {
        if (threadIdx.x == 0) counts[k] = total;
        return;
    }
    const uint64_t before_end = mode == 0 ? block_scan_u64(mine_end, &total) : 0;
    uint64_t o = offs[k] + before, oe = offs[k] + before_end;
    for (uint32_t q = q0; q < q1; ++q) {
        if (mode == 0) {
            uint64_t s, e;
            word_edges(w, q, s, e);
            while (s) { out_a[o++] = q * 64 + (uint32_t)__ffsll((long long)s) - 1; s &= s - 1; }
            while (e) { out_b[oe++] = q * 64 + (uint32_t)__ffsll((long long)e) - 1; e &= e - 1; }
        } else {
            uint64_t x = w[q];
            while (x) { out_a[o++] = q * 64 + (uint32_t)__ffsll((long long)x) - 1; x &= x - 1; }
        }
    }
    if (mode == 0 && tail_end && (w[nw - 1] >> 63)) out_b[oe++] = universe;
}

// Copy scratch rows into their bitmap slots (kNoBitmap: none).
__global__ void copy_rows_kernel(const uint64_t* __restrict__ scratch, uint32_t ld, uint32_t row0,
                                 uint32_t rows, const uint32_t* __restrict__ bm_row,
                                 uint64_t* __restrict__ store) {
    const uint32_t k = blockIdx.y;
    if (k >= rows) return;
    const uint32_t slot = bm_row[row0 + k];
    if (slot == kNoBitmap) return;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < ld; q += gridDim.x * blockDim.x)
        store[(uint64_t)slot * ld + q] = scratch[(uint64_t)k * ld + q];
}

// ---- RowMeta of runs-only and bitmap-only rows ---------------------------------

// meta (8 u32): cardinality, n_runs, n_words, n_nonzero_w, first_set,
// last_set (empty row: 0xFFFFFFFF, 0), has_rank (mean run length >= 512),
// has_occ (cardinality * 512 > universe): build_row's fields and gates.
__device__ inline void write_meta(uint32_t* m, uint64_t card, uint64_t runs, uint32_t nw,
                                  uint64_t nnz, uint32_t first, uint32_t last, uint32_t universe) {
    m[0] = (uint32_t)card; m[1] = (uint32_t)runs; m[2] = nw; m[3] = (uint32_t)nnz;
    m[4] = card ? first : 0xFFFFFFFFu;
    m[5] = card ? last : 0;
    const uint64_t mean_rl = runs ? card / runs : 0;
    m[6] = mean_rl >= 512;
    m[7] = card * 512 > (uint64_t)universe;
}

__global__ __launch_bounds__(256) void meta_from_runs_kernel(
    const uint64_t* __restrict__ run_off, const uint32_t* __restrict__ rs,
    const uint32_t* __restrict__ re, uint32_t n, uint32_t nw, uint32_t universe,
    uint32_t* __restrict__ meta) {
    const uint32_t r = blockIdx.x;
    if (r >= n) return;
    const uint64_t a = run_off[r], b = run_off[r + 1];
    uint64_t card = 0, nnz = 0;
    for (uint64_t k = a + threadIdx.x; k < b; k += blockDim.x) {
        card += re[k] - rs[k];
        const uint32_t ws = rs[k] >> 6, we = (re[k] - 1) >> 6;
        nnz += we - ws + 1;
        if (k > a && ((re[k - 1] - 1) >> 6) == ws) nnz -= 1;   // shares a word with the previous run
    }
    card = block_sum_u64(card);
    nnz = block_sum_u64(nnz);
    if (threadIdx.x == 0)
        write_meta(meta + (uint64_t)r * 8, card, b - a, nw, nnz, b > a ? rs[a] : 0,
                   b > a ? re[b - 1] - 1 : 0, universe);
}

__global__ __launch_bounds__(256) void meta_from_bitmap_kernel(
    const uint64_t* __restrict__ store, uint32_t ld, const uint32_t* __restrict__ bm_row,
    uint32_t n, uint32_t nw, uint32_t universe, uint32_t* __restrict__ meta) {
    const uint32_t r = blockIdx.x;
    if (r >= n) return;
    const uint64_t* w = store + (uint64_t)bm_row[r] * ld;
    uint64_t card = 0, runs = 0, nnz = 0;
    uint32_t first = 0xFFFFFFFFu, last = 0;
    for (uint32_t q = threadIdx.x; q < nw; q += blockDim.x) {
        const uint64_t x = w[q];
        if (!x) continue;
        uint64_t s, e;
        word_edges(w, q, s, e);
        card += __popcll(x);
        runs += __popcll(s);
        nnz += 1;
        first = min(first, q * 64 + (uint32_t)__ffsll((long long)x) - 1);
        last = max(last, q * 64 + 63 - (uint32_t)__clzll((long long)x));
    }
    card = block_sum_u64(card);
    runs = block_sum_u64(runs);
    nnz = block_sum_u64(nnz);
    __shared__ uint32_t sf[256], sl[256];
    sf[threadIdx.x] = first;
    sl[threadIdx.x] = last;
    __syncthreads();
    for (uint32_t st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            sf[threadIdx.x] = min(sf[threadIdx.x], sf[threadIdx.x + st]);
            sl[threadIdx.x] = max(sl[threadIdx.x], sl[threadIdx.x + st]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        write_meta(meta + (uint64_t)r * 8, card, runs, nw, nnz, sf[0], sl[0], universe);
}

// ---- rank9 and zone map ---------------------------------------------------------

// build_rank (kernels/storm_repr.cpp:25-45): per block b of 8 words,
// rank[2b] = set bits before the block, rank[2b+1] = 9-bit counts of words
// [8b, 8b+k) for k = 1..7 at bits 9(k-1); a sentinel block holds the total.
__global__ __launch_bounds__(256) void build_rank_kernel(
    const uint64_t* __restrict__ store, uint32_t ld, const uint32_t* __restrict__ bm_row,
    const uint32_t* __restrict__ rank_row, uint32_t n, uint32_t nw, uint32_t rank_ld,
    uint64_t* __restrict__ rank) {
    const uint32_t r = blockIdx.x;
    if (r >= n || rank_row[r] == kNoBitmap) return;
    const uint64_t* w = store + (uint64_t)bm_row[r] * ld;
    uint64_t* out = rank + (uint64_t)rank_row[r] * rank_ld;
    const uint32_t nblk = (nw + 7) / 8;
    uint64_t carry = 0;
    for (uint32_t b0 = 0; b0 < nblk; b0 += blockDim.x) {
        const uint32_t b = b0 + threadIdx.x;
        uint64_t rel = 0, sub = 0;
        if (b < nblk) {
            for (uint32_t k = 0; k < 8; ++k) {
                if (k >= 1) sub |= (rel & 0x1FF) << (9 * (k - 1));
                const uint32_t wi = b * 8 + k;
                if (wi < nw) rel += __popcll(w[wi]);
            }
        }
        uint64_t total;
        const uint64_t before = block_scan_u64(rel, &total);
        if (b < nblk) {
            out[2 * (uint64_t)b] = carry + before;
            out[2 * (uint64_t)b + 1] = sub;
        }
        carry += total;
    }
    if (threadIdx.x == 0) {
        out[2 * (uint64_t)nblk] = carry;
        out[2 * (uint64_t)nblk + 1] = 0;
    }
}

// build_occ (kernels/storm_repr.cpp:47-59): bit b set when any word of
// [8b, min(8b + 8, nw)) is non-zero. One lane per bin, 64 bins per word.
__global__ __launch_bounds__(256) void build_occ_kernel(
    const uint64_t* __restrict__ store, uint32_t ld, const uint32_t* __restrict__ bm_row,
    const uint32_t* __restrict__ occ_row, uint32_t n, uint32_t nw, uint32_t occ_ld,
    uint64_t* __restrict__ occ) {
    const uint32_t r = blockIdx.y;
    if (r >= n || occ_row[r] == kNoBitmap) return;
    const uint64_t* w = store + (uint64_t)bm_row[r] * ld;
    uint64_t* out = occ + (uint64_t)occ_row[r] * occ_ld;
    const uint32_t bins = (nw + 7) / 8;
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < occ_ld * 64; b0 += gridDim.x * blockDim.x) {
        const uint32_t b = b0 + threadIdx.x;
        bool any = false;
        if (b < bins)
            for (uint32_t k = b * 8; k < min(b * 8 + 8, nw); ++k) any |= w[k] != 0;
        const uint64_t m = __ballot(any);
        if ((threadIdx.x & 63) == 0 && b / 64 < occ_ld) out[b / 64] = m;
    }
}

// ---- column closed forms --------------------------------------------------------

__global__ void col_from_lists_kernel(const uint32_t* __restrict__ list, uint64_t total,
                                      uint32_t* __restrict__ col) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(&col[list[i]], 1u);
}

// Runs add +1 at their start and -1 at their end into a difference array
// (a prefix sum then gives the column counts).
__global__ void col_from_runs_kernel(const uint32_t* __restrict__ rs,
                                     const uint32_t* __restrict__ re, uint64_t total,
                                     int* __restrict__ diff) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        atomicAdd(&diff[rs[i]], 1);
        atomicAdd(&diff[re[i]], -1);
    }
}

// Prefix sums over a u32 / i32 array in three phases: per-chunk sums, a scan
// of the chunk sums (one workgroup), then each chunk rescanned from its
// offset. kChunk elements per workgroup (16 per thread).
constexpr uint32_t kScanChunk = 256 * 16;

template <class T>
__global__ __launch_bounds__(256) void chunk_sums_kernel(const T* __restrict__ x, uint32_t n,
                                                         int64_t* __restrict__ sums) {
    const uint64_t c0 = (uint64_t)blockIdx.x * kScanChunk;
    int64_t v = 0;
    for (uint32_t k = 0; k < 16; ++k) {
        const uint64_t b = c0 + threadIdx.x * 16 + k;
        if (b < n) v += (int64_t)x[b];
    }
    const uint64_t t = block_sum_u64((uint64_t)v);
    if (threadIdx.x == 0) sums[blockIdx.x] = (int64_t)t;
}

__global__ __launch_bounds__(256) void scan_sums_kernel(int64_t* __restrict__ sums, uint32_t m) {
    uint64_t carry = 0;
    for (uint32_t b0 = 0; b0 < m; b0 += blockDim.x) {
        const uint32_t b = b0 + threadIdx.x;
        const uint64_t v = b < m ? (uint64_t)sums[b] : 0;
        uint64_t total;
        const uint64_t before = block_scan_u64(v, &total);
        if (b < m) sums[b] = (int64_t)(carry + before);
        carry += total;
    }
}

// Inclusive prefix in place (out == x as int), or exclusive into out64 (with
// out64[n] = total) when out64 is set.
template <class T>
__global__ __launch_bounds__(256) void chunk_scan_kernel(T* __restrict__ x, uint32_t n,
                                                         const int64_t* __restrict__ sums,
                                                         unsigned long long* __restrict__ out64) {
    const uint64_t c0 = (uint64_t)blockIdx.x * kScanChunk;
    int64_t v[16], mine = 0;
    for (uint32_t k = 0; k < 16; ++k) {
        const uint64_t b = c0 + threadIdx.x * 16 + k;
        v[k] = b < n ? (int64_t)x[b] : 0;
        mine += v[k];
    }
    uint64_t total;
    int64_t acc = sums[blockIdx.x] + (int64_t)block_scan_u64((uint64_t)mine, &total);
    for (uint32_t k = 0; k < 16; ++k) {
        const uint64_t b = c0 + threadIdx.x * 16 + k;
        if (b >= n) break;
        if (out64) out64[b] = (unsigned long long)acc;
        acc += v[k];
        if (!out64) x[b] = (T)acc;
        if (out64 && b + 1 == n) out64[n] = (unsigned long long)acc;
    }
}

// Bitmap rows: thread per word column, a 13-bit vertical counter per bit
// position (bit-sliced ripple-carry adds), then the 64 counts added out.
__global__ void col_from_bitmaps_kernel(const uint64_t* __restrict__ store, uint32_t ld,
                                        const uint32_t* __restrict__ bm_row, uint32_t n,
                                        uint32_t nw, uint32_t* __restrict__ col) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nw) return;
    constexpr int kBits = 13;   // counts up to 8191 per pass
    for (uint32_t r0 = 0; r0 < n; r0 += (1u << kBits) - 1) {
        uint64_t c[kBits] = {};
        const uint32_t r1 = min(n, r0 + (1u << kBits) - 1);
        for (uint32_t r = r0; r < r1; ++r) {
            const uint32_t slot = bm_row[r];
            if (slot == kNoBitmap) continue;
            uint64_t carry = store[(uint64_t)slot * ld + q];
#pragma unroll
            for (int k = 0; k < kBits; ++k) {
                const uint64_t t = c[k] & carry;
                c[k] ^= carry;
                carry = t;
            }
        }
        for (int b = 0; b < 64; ++b) {
            uint32_t v = 0;
#pragma unroll
            for (int k = 0; k < kBits; ++k) v |= (uint32_t)((c[k] >> b) & 1) << k;
            if (v) col[(uint64_t)q * 64 + b] += v;
        }
    }
}

__global__ void pairs_total_kernel(const uint32_t* __restrict__ col, uint32_t n,
                                   unsigned long long* __restrict__ total) {
    uint64_t s = 0;
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < n; b += gridDim.x * blockDim.x) {
        const uint64_t c = col[b];
        s += c * (c ? c - 1 : 0) / 2;
    }
    s = wave_sum_u64(s);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(total, (unsigned long long)s);
}

// Row sums from the row's authoritative representation (list, else runs,
// else bitmap). For runs, pre[x] = sum_{b < x} c_b (exclusive prefix).
__global__ __launch_bounds__(256) void row_sums_kernel(DevCorpus D, int src,
                                                       const uint32_t* __restrict__ col,
                                                       const unsigned long long* __restrict__ pre,
                                                       unsigned long long* __restrict__ out) {
    const uint32_t r = blockIdx.x;
    if (r >= D.n) return;
    uint64_t s = 0;
    if (src == 0) {
        for (uint64_t i = D.list_off[r] + threadIdx.x; i < D.list_off[r + 1]; i += blockDim.x)
            s += col[D.list[i]] - 1;
    } else if (src == 1) {
        for (uint64_t i = D.run_off[r] + threadIdx.x; i < D.run_off[r + 1]; i += blockDim.x)
            s += (pre[D.run_end[i]] - pre[D.run_start[i]]) - (D.run_end[i] - D.run_start[i]);
    } else {
        const uint64_t* w = D.bitmap + (uint64_t)D.bm_row[r] * D.ld;
        for (uint32_t q = threadIdx.x; q < D.nw; q += blockDim.x) {
            uint64_t x = w[q];
            while (x) {
                s += col[(uint64_t)q * 64 + __ffsll((long long)x) - 1] - 1;
                x &= x - 1;
            }
        }
    }
    s = block_sum_u64(s);
    if (threadIdx.x == 0) out[r] = s;
}

struct Dev {
    void* p = nullptr;
    ~Dev() { if (p) (void)hipFree(p); }
};

uint32_t grid_for(uint64_t n, uint32_t block = 256, uint32_t cap = 65535u * 4) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + block - 1) / block, cap));
}

// In-place inclusive prefix of x[0, n) (out64 null), or the exclusive prefix
// of x into out64[0, n] (out64[n] the total).
template <class T>
int prefix_scan(hipStream_t s, T* x, uint32_t n, unsigned long long* out64) {
    const uint32_t m = std::max<uint32_t>(1, (uint32_t)(((uint64_t)n + kScanChunk - 1) / kScanChunk));
    Dev sums;
    SH_HIP(hipMalloc(&sums.p, (size_t)m * 8));
    int64_t* d = (int64_t*)sums.p;
    hipLaunchKernelGGL(chunk_sums_kernel<T>, dim3(m), dim3(256), 0, s, x, n, d);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(256), 0, s, d, m);
    hipLaunchKernelGGL(chunk_scan_kernel<T>, dim3(m), dim3(256), 0, s, x, n, (const int64_t*)d,
                       out64);
    SH_HIP(hipGetLastError());
    SH_HIP(hipStreamSynchronize(s));   // the chunk sums are freed on return
    return STORM_HIP_OK;
}

// Scatter runs into bitmap slots (rows whose bm_row is set).
__global__ void runs_to_bits_kernel(const uint64_t* __restrict__ run_off,
                                    const uint32_t* __restrict__ rs,
                                    const uint32_t* __restrict__ re, uint32_t n,
                                    const uint32_t* __restrict__ bm_row, uint32_t ld,
                                    unsigned long long* __restrict__ store) {
    const uint32_t r = blockIdx.x;
    if (r >= n || bm_row[r] == kNoBitmap) return;
    unsigned long long* w = store + (uint64_t)bm_row[r] * ld;
    for (uint64_t k = run_off[r] + threadIdx.x; k < run_off[r + 1]; k += blockDim.x) {
        const uint32_t s = rs[k], e = re[k];
        const uint32_t ws = s >> 6, we = (e - 1) >> 6;
        const uint64_t lo = ~0ull << (s & 63), hi = ~0ull >> (63 - ((e - 1) & 63));
        if (ws == we) { atomicOr(&w[ws], lo & hi); continue; }
        atomicOr(&w[ws], lo);
        for (uint32_t q = ws + 1; q < we; ++q) w[q] = ~0ull;   // whole words belong to this run
        atomicOr(&w[we], hi);
    }
}

}  // namespace

// ---- host side -------------------------------------------------------------------

int fill_bitmaps_from_runs(storm_hip_corpus* c, hipStream_t s) {
    if (!c->n_rows) return STORM_HIP_OK;
    hipLaunchKernelGGL(runs_to_bits_kernel, dim3(c->n_rows), dim3(256), 0, s, c->d_run_off,
                       c->d_run_start, c->d_run_end, c->n_rows, c->d_bm_row, c->ld,
                       (unsigned long long*)c->d_bitmap);
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

int build_indexes(storm_hip_corpus* c, hipStream_t s) {
    // Rows holding a bitmap get rank9 / a zone map exactly when build_row's
    // gates (meta fields 6 and 7) say so.
    for (void** p : {(void**)&c->d_rank, (void**)&c->d_rank_row, (void**)&c->d_occ,
                     (void**)&c->d_occ_row})
        if (*p) { SH_HIP(hipFree(*p)); *p = nullptr; }
    c->device_bytes -= c->index_bytes;
    c->index_bytes = 0;
    const uint32_t n = c->n_rows, nw = c->n_words;
    c->rank_ld = 2 * ((nw + 7) / 8 + 1);
    c->occ_ld = std::max(1u, ((nw + 7) / 8 + 63) / 64);
    std::vector<uint32_t> rr((size_t)n + 1, kNoBitmap), orow((size_t)n + 1, kNoBitmap);
    uint32_t nr = 0, no = 0;
    for (uint32_t r = 0; r < n; ++r) {
        if (c->h_bm_row[r] == kNoBitmap) continue;
        if (c->h_meta[(size_t)r * 8 + 6]) rr[r] = nr++;
        if (c->h_meta[(size_t)r * 8 + 7]) orow[r] = no++;
    }
    c->n_rank_rows = nr;
    c->n_occ_rows = no;
    SH_TRY(upload_vec(c->d_rank_row, rr, s, c->index_bytes));
    SH_TRY(upload_vec(c->d_occ_row, orow, s, c->index_bytes));
    const uint64_t rb = std::max<uint64_t>(1, (uint64_t)nr * c->rank_ld) * 8;
    const uint64_t ob = std::max<uint64_t>(1, (uint64_t)no * c->occ_ld) * 8;
    SH_HIP(hipMalloc(&c->d_rank, rb));
    SH_HIP(hipMalloc(&c->d_occ, ob));
    c->index_bytes += rb + ob;
    c->device_bytes += c->index_bytes;
    if (nr) {
        hipLaunchKernelGGL(build_rank_kernel, dim3(n), dim3(256), 0, s, c->d_bitmap, c->ld,
                           c->d_bm_row, c->d_rank_row, n, nw, c->rank_ld, c->d_rank);
        SH_HIP(hipGetLastError());
    }
    if (no) {
        const dim3 grid(std::min<uint32_t>((c->occ_ld * 64 + 255) / 256, 64), n);
        hipLaunchKernelGGL(build_occ_kernel, grid, dim3(256), 0, s, c->d_bitmap, c->ld,
                           c->d_bm_row, c->d_occ_row, n, nw, c->occ_ld, c->d_occ);
        SH_HIP(hipGetLastError());
    }
    return STORM_HIP_OK;
}

namespace {

// Rows of a runs-only or list-producing generator go through scratch bitmaps,
// in chunks of rows whose scratch stays within this many bytes.
constexpr uint64_t kScratchBytes = 8ull << 30;

// Meta of a corpus whose runs are in place, then the host copy.
int meta_from_runs(storm_hip_corpus* c, hipStream_t s) {
    std::vector<uint32_t> zero((size_t)c->n_rows * 8, 0);
    SH_TRY(upload_vec(c->d_meta, zero, s, c->device_bytes));
    if (c->n_rows) {
        hipLaunchKernelGGL(meta_from_runs_kernel, dim3(c->n_rows), dim3(256), 0, s, c->d_run_off,
                           c->d_run_start, c->d_run_end, c->n_rows, c->n_words, c->universe,
                           c->d_meta);
        SH_HIP(hipGetLastError());
    }
    c->h_meta.assign((size_t)c->n_rows * 8, 0);
    if (c->n_rows)
        SH_HIP(hipMemcpyAsync(c->h_meta.data(), c->d_meta, c->h_meta.size() * 4,
                              hipMemcpyDeviceToHost, s));
    SH_HIP(hipStreamSynchronize(s));
    return STORM_HIP_OK;
}

// Empty arenas (n + 1 zero offsets) for the representations a corpus lacks.
int empty_offsets(uint64_t*& d, uint32_t n, hipStream_t s, uint64_t& bytes) {
    std::vector<uint64_t> z((size_t)n + 1, 0);
    return upload_vec(d, z, s, bytes);
}

int gen_runs_corpus(storm_hip_ctx* ctx, storm_hip_corpus* c, double lambda, double mean_run,
                    double frac, uint64_t seed, uint32_t flags) {
    hipStream_t s = ctx->stream;
    const uint32_t n = c->n_rows, ld = c->ld;
    const uint64_t row_bytes = (uint64_t)ld * 8;
    const uint32_t chunk = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(n, kScratchBytes / row_bytes));
    Dev scratch, cnt, offs;
    SH_HIP(hipMalloc(&scratch.p, (uint64_t)chunk * row_bytes));
    SH_HIP(hipMalloc(&cnt.p, ((size_t)n + 1) * 8));
    // Pass 1: run counts of every row (rows regenerated per pass).
    std::vector<uint64_t> h_cnt(n, 0);
    uint64_t* d_cnt = (uint64_t*)cnt.p;
    uint64_t* d_scr = (uint64_t*)scratch.p;
    auto gen_chunk = [&](uint32_t r0, uint32_t rows) -> int {
        hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_for((uint64_t)rows * ld)), dim3(256), 0, s,
                           d_scr, (uint64_t)rows * ld);
        hipLaunchKernelGGL(gen_runs_kernel, dim3((rows + 63) / 64), dim3(64), 0, s, d_scr, ld, r0,
                           rows, c->universe, lambda, mean_run, frac, seed);
        SH_HIP(hipGetLastError());
        return STORM_HIP_OK;
    };
    for (uint32_t r0 = 0; r0 < n; r0 += chunk) {
        const uint32_t rows = std::min(chunk, n - r0);
        SH_TRY(gen_chunk(r0, rows));
        hipLaunchKernelGGL(extract_kernel, dim3(rows), dim3(256), 0, s, d_scr, ld, c->n_words,
                           c->universe, rows, 0, 0, d_cnt + r0, nullptr, nullptr, nullptr);
        SH_HIP(hipGetLastError());
    }
    if (n) SH_HIP(hipMemcpyAsync(h_cnt.data(), d_cnt, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    SH_HIP(hipStreamSynchronize(s));
    std::vector<uint64_t> off((size_t)n + 1, 0);
    for (uint32_t r = 0; r < n; ++r) off[r + 1] = off[r] + h_cnt[r];
    SH_TRY(upload_vec(c->d_run_off, off, s, c->device_bytes));
    const uint64_t total = std::max<uint64_t>(off[n], 1);
    SH_HIP(hipMalloc((void**)&c->d_run_start, total * 4));
    SH_HIP(hipMalloc((void**)&c->d_run_end, total * 4));
    c->device_bytes += total * 8;
    // Placement of bitmaps needs the cardinalities: count them from the runs
    // after pass 2; the scratch rows are copied when the store exists, so
    // rows are regenerated a third time only for the bitmap rows' chunks.
    for (uint32_t r0 = 0; r0 < n; r0 += chunk) {
        const uint32_t rows = std::min(chunk, n - r0);
        SH_TRY(gen_chunk(r0, rows));
        hipLaunchKernelGGL(extract_kernel, dim3(rows), dim3(256), 0, s, d_scr, ld, c->n_words,
                           c->universe, rows, 0, 1, nullptr, c->d_run_off + r0, c->d_run_start,
                           c->d_run_end);
        SH_HIP(hipGetLastError());
    }
    c->has_runs = true;
    c->has_lists = false;
    c->has_ewah = false;
    c->has_comp = false;
    SH_TRY(empty_offsets(c->d_list_off, n, s, c->device_bytes));
    SH_TRY(empty_offsets(c->d_ewah_off, n, s, c->device_bytes));
    SH_TRY(empty_offsets(c->d_mk_off, n, s, c->device_bytes));
    SH_TRY(empty_offsets(c->d_comp_off, n, s, c->device_bytes));
    SH_TRY(meta_from_runs(c, s));
    (void)flags;
    return STORM_HIP_OK;
}

// Sparse Bernoulli rows as lists (two passes through scratch bitmaps).
int gen_skip_lists(storm_hip_ctx* ctx, storm_hip_corpus* c, double d, uint64_t seed) {
    hipStream_t s = ctx->stream;
    const uint32_t n = c->n_rows, ld = c->ld;
    const uint64_t row_bytes = (uint64_t)ld * 8;
    const uint32_t chunk = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(n, kScratchBytes / row_bytes));
    Dev scratch, cnt;
    SH_HIP(hipMalloc(&scratch.p, (uint64_t)chunk * row_bytes));
    SH_HIP(hipMalloc(&cnt.p, ((size_t)n + 1) * 8));
    uint64_t* d_cnt = (uint64_t*)cnt.p;
    uint64_t* d_scr = (uint64_t*)scratch.p;
    auto gen_chunk = [&](uint32_t r0, uint32_t rows) -> int {
        hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_for((uint64_t)rows * ld)), dim3(256), 0, s,
                           d_scr, (uint64_t)rows * ld);
        hipLaunchKernelGGL(gen_skip_kernel, dim3((rows + 63) / 64), dim3(64), 0, s, d_scr, ld, r0,
                           rows, c->universe, d, seed);
        SH_HIP(hipGetLastError());
        return STORM_HIP_OK;
    };
    for (uint32_t r0 = 0; r0 < n; r0 += chunk) {
        const uint32_t rows = std::min(chunk, n - r0);
        SH_TRY(gen_chunk(r0, rows));
        hipLaunchKernelGGL(extract_kernel, dim3(rows), dim3(256), 0, s, d_scr, ld, c->n_words,
                           c->universe, rows, 1, 0, d_cnt + r0, nullptr, nullptr, nullptr);
        SH_HIP(hipGetLastError());
    }
    std::vector<uint64_t> h_cnt(n, 0);
    if (n) SH_HIP(hipMemcpyAsync(h_cnt.data(), d_cnt, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    SH_HIP(hipStreamSynchronize(s));
    std::vector<uint64_t> off((size_t)n + 1, 0);
    for (uint32_t r = 0; r < n; ++r) off[r + 1] = off[r] + h_cnt[r];
    SH_TRY(upload_vec(c->d_list_off, off, s, c->device_bytes));
    const uint64_t total = std::max<uint64_t>(off[n], 1);
    SH_HIP(hipMalloc((void**)&c->d_list, total * 4));
    c->device_bytes += total * 4;
    for (uint32_t r0 = 0; r0 < n; r0 += chunk) {
        const uint32_t rows = std::min(chunk, n - r0);
        SH_TRY(gen_chunk(r0, rows));
        hipLaunchKernelGGL(extract_kernel, dim3(rows), dim3(256), 0, s, d_scr, ld, c->n_words,
                           c->universe, rows, 1, 1, nullptr, c->d_list_off + r0, c->d_list,
                           nullptr);
        SH_HIP(hipGetLastError());
    }
    SH_HIP(hipStreamSynchronize(s));
    return STORM_HIP_OK;
}

// Dense Bernoulli rows: every row a bitmap, nothing else.
int gen_bits_corpus(storm_hip_ctx* ctx, storm_hip_corpus* c, double p, uint64_t seed) {
    hipStream_t s = ctx->stream;
    const uint32_t n = c->n_rows, ld = c->ld;
    c->has_lists = c->has_runs = c->has_ewah = c->has_comp = false;
    SH_TRY(empty_offsets(c->d_list_off, n, s, c->device_bytes));
    SH_TRY(empty_offsets(c->d_run_off, n, s, c->device_bytes));
    SH_TRY(empty_offsets(c->d_ewah_off, n, s, c->device_bytes));
    SH_TRY(empty_offsets(c->d_mk_off, n, s, c->device_bytes));
    SH_TRY(empty_offsets(c->d_comp_off, n, s, c->device_bytes));
    // The store: one slot per row (slot = row) plus the zero row.
    c->h_bm_row.assign((size_t)n + 1, 0);
    std::iota(c->h_bm_row.begin(), c->h_bm_row.end(), 0u);
    c->n_bitmap_rows = n;
    c->bm_capacity = n;
    c->bitmaps_complete = true;
    SH_TRY(upload_vec(c->d_bm_row, c->h_bm_row, s, c->device_bytes));
    const uint64_t bytes = ((uint64_t)n + 1) * ld * 8;
    SH_HIP(hipMalloc((void**)&c->d_bitmap, bytes));
    c->device_bytes += bytes;
    hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_for((uint64_t)ld)), dim3(256), 0, s,
                       c->d_bitmap + (uint64_t)n * ld, (uint64_t)ld);
    // ceil(p * 2^53): (next() >> 11) < ceil(p53) exactly when (double)(next() >> 11) < p53.
    const double p53 = p * 9007199254740992.0;
    uint64_t thr = p53 <= 0 ? 0 : (p53 >= 9007199254740992.0 ? (1ull << 53) : (uint64_t)p53);
    if (p53 > 0 && p53 < 9007199254740992.0 && (double)thr < p53) ++thr;
    const uint64_t seed0 = seed ? seed : kGolden;
    const uint32_t rows_per = (uint32_t)std::max<uint64_t>(1, (1ull << 28) / ld);
    for (uint32_t r0 = 0; r0 < n; r0 += rows_per) {
        const uint32_t rows = std::min(rows_per, n - r0);
        hipLaunchKernelGGL(gen_bits_kernel, dim3((uint32_t)(((uint64_t)rows * ld + 255) / 256)),
                           dim3(256), 0, s, c->d_bitmap, ld, r0, r0, rows, c->universe, thr, seed0);
        SH_HIP(hipGetLastError());
    }
    std::vector<uint32_t> zero((size_t)n * 8, 0);
    SH_TRY(upload_vec(c->d_meta, zero, s, c->device_bytes));
    if (n) {
        hipLaunchKernelGGL(meta_from_bitmap_kernel, dim3(n), dim3(256), 0, s, c->d_bitmap, ld,
                           c->d_bm_row, n, c->n_words, c->universe, c->d_meta);
        SH_HIP(hipGetLastError());
    }
    c->h_meta.assign((size_t)n * 8, 0);
    if (n) SH_HIP(hipMemcpyAsync(c->h_meta.data(), c->d_meta, c->h_meta.size() * 4,
                                 hipMemcpyDeviceToHost, s));
    SH_HIP(hipStreamSynchronize(s));
    return STORM_HIP_OK;
}

}  // namespace

}  // namespace stormhip

using namespace stormhip;

extern "C" {

int storm_hip_corpus_gen(storm_hip_ctx* ctx, uint32_t kind, uint32_t n_rows, uint32_t universe,
                         double p1, double p2, double p3, uint64_t seed, uint32_t flags,
                         storm_hip_corpus** out) {
    if (!ctx || !out) return STORM_HIP_ERR_INVALID;
    *out = nullptr;
    if (universe == 0 || (flags & 3u) == 3u) return STORM_HIP_ERR_INVALID;
    if (kind == STORM_HIP_GEN_RUNS && !(p1 >= 0.0 && p2 >= 0.0 && p3 >= 0.0)) return STORM_HIP_ERR_INVALID;
    if (kind == STORM_HIP_GEN_BERNOULLI && !(p1 >= 0.0 && p1 <= 1.0)) return STORM_HIP_ERR_INVALID;
    if (kind != STORM_HIP_GEN_RUNS && kind != STORM_HIP_GEN_BERNOULLI) {
        set_error("device generation takes the Bernoulli and run-heavy workloads");
        return STORM_HIP_ERR_UNSUPPORTED;
    }
    SH_HIP(hipSetDevice(ctx->device));
    storm_hip_corpus* c = corpus_new(ctx, n_rows, universe, flags);
    if (!c) return STORM_HIP_ERR_NOMEM;
    auto fail = [&](int st) { storm_hip_corpus_free(c); return st; };
    int st = STORM_HIP_OK;
    if (kind == STORM_HIP_GEN_BERNOULLI && p1 * 32.0 >= 1.0) {
        // A bitmap is no larger than the list: bitmaps only.
        st = gen_bits_corpus(ctx, c, p1, seed);
        if (st == STORM_HIP_OK) st = corpus_finish_store(ctx, c);
    } else if (kind == STORM_HIP_GEN_BERNOULLI) {
        st = gen_skip_lists(ctx, c, p1, seed);
        if (st == STORM_HIP_OK) st = device_build(ctx, c, flags);
        if (st == STORM_HIP_OK)
            st = corpus_finish(ctx, c, bitmap_rows(c->h_meta, n_rows, universe, flags & 3u));
    } else {
        st = gen_runs_corpus(ctx, c, p1, p2, p3, seed, flags);
        if (st == STORM_HIP_OK)
            st = corpus_finish(ctx, c, bitmap_rows(c->h_meta, n_rows, universe, flags & 3u));
    }
    if (st != STORM_HIP_OK) return fail(st);
    *out = c;
    return STORM_HIP_OK;
}

int storm_hip_corpus_closed_form(storm_hip_ctx* ctx, const storm_hip_corpus* c, uint64_t* total,
                                 uint64_t* row_sums) {
    if (!ctx || !c || c->ctx != ctx || !total) return STORM_HIP_ERR_INVALID;
    SH_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    *total = 0;
    const uint32_t n = c->n_rows, U = c->dense ? c->n_words * 64 : c->universe;
    if (n == 0) return STORM_HIP_OK;
    Dev col, pre, tot, rs;
    SH_HIP(hipMalloc(&col.p, ((uint64_t)U + 1) * 4));
    SH_HIP(hipMalloc(&tot.p, 8));
    hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_for(((uint64_t)U + 2) / 2)), dim3(256), 0, s,
                       (uint64_t*)col.p, ((uint64_t)U + 1) / 2);
    if ((U + 1) % 2) SH_HIP(hipMemsetAsync((uint32_t*)col.p + U, 0, 4, s));
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, (uint64_t*)tot.p, 1);
    // Source of truth per corpus: lists, else runs, else bitmaps.
    int src = 2;
    DevCorpus D{};
    D.n = n; D.universe = U; D.nw = c->n_words; D.ld = c->ld;
    if (c->dense) {
        std::vector<uint32_t> ident((size_t)n + 1);
        std::iota(ident.begin(), ident.end(), 0u);
        uint64_t b = 0;
        uint32_t* d_ident = nullptr;
        SH_TRY(upload_vec(d_ident, ident, s, b));
        Dev guard;
        guard.p = d_ident;
        hipLaunchKernelGGL(col_from_bitmaps_kernel, dim3(grid_for(c->n_words)), dim3(256), 0, s,
                           c->d_words, c->ld, d_ident, n, c->n_words, (uint32_t*)col.p);
        SH_HIP(hipGetLastError());
        D.bitmap = c->d_words; D.bm_row = d_ident;
        hipLaunchKernelGGL(pairs_total_kernel, dim3(grid_for(U, 256, 4096)), dim3(256), 0, s,
                           (const uint32_t*)col.p, U, (unsigned long long*)tot.p);
        if (row_sums) {
            SH_HIP(hipMalloc(&rs.p, (size_t)n * 8));
            hipLaunchKernelGGL(row_sums_kernel, dim3(n), dim3(256), 0, s, D, 2,
                               (const uint32_t*)col.p, nullptr, (unsigned long long*)rs.p);
        }
        SH_HIP(hipGetLastError());
        SH_HIP(hipMemcpyAsync(total, tot.p, 8, hipMemcpyDeviceToHost, s));
        if (row_sums) SH_HIP(hipMemcpyAsync(row_sums, rs.p, (size_t)n * 8, hipMemcpyDeviceToHost, s));
        SH_HIP(hipStreamSynchronize(s));
        return STORM_HIP_OK;
    }
    D.list_off = c->d_list_off; D.list = c->d_list;
    D.run_off = c->d_run_off; D.run_start = c->d_run_start; D.run_end = c->d_run_end;
    D.bm_row = c->d_bm_row; D.bitmap = c->d_bitmap;
    uint64_t list_total = 0, run_total = 0;
    SH_HIP(hipMemcpy(&list_total, c->d_list_off + n, 8, hipMemcpyDeviceToHost));
    if (c->has_runs) SH_HIP(hipMemcpy(&run_total, c->d_run_off + n, 8, hipMemcpyDeviceToHost));
    if (c->has_lists) {
        src = 0;
        if (list_total)
            hipLaunchKernelGGL(col_from_lists_kernel, dim3(grid_for(list_total, 256, 8192)),
                               dim3(256), 0, s, c->d_list, list_total, (uint32_t*)col.p);
    } else if (c->has_runs) {
        src = 1;
        if (run_total)
            hipLaunchKernelGGL(col_from_runs_kernel, dim3(grid_for(run_total, 256, 8192)),
                               dim3(256), 0, s, c->d_run_start, c->d_run_end, run_total,
                               (int*)col.p);
        SH_TRY(prefix_scan<int>(s, (int*)col.p, U, nullptr));
    } else {
        if (!c->bitmaps_complete) {
            set_error("closed form: rows without a list, runs or bitmap");
            return STORM_HIP_ERR_INVALID;
        }
        hipLaunchKernelGGL(col_from_bitmaps_kernel, dim3(grid_for(c->n_words)), dim3(256), 0, s,
                           c->d_bitmap, c->ld, c->d_bm_row, n, c->n_words, (uint32_t*)col.p);
    }
    SH_HIP(hipGetLastError());
    hipLaunchKernelGGL(pairs_total_kernel, dim3(grid_for(U, 256, 4096)), dim3(256), 0, s,
                       (const uint32_t*)col.p, U, (unsigned long long*)tot.p);
    SH_HIP(hipGetLastError());
    if (row_sums) {
        SH_HIP(hipMalloc(&rs.p, (size_t)n * 8));
        if (src == 1) {
            SH_HIP(hipMalloc(&pre.p, ((uint64_t)U + 1) * 8));
            SH_TRY(prefix_scan<uint32_t>(s, (uint32_t*)col.p, U, (unsigned long long*)pre.p));
        }
        hipLaunchKernelGGL(row_sums_kernel, dim3(n), dim3(256), 0, s, D, src,
                           (const uint32_t*)col.p, (const unsigned long long*)pre.p,
                           (unsigned long long*)rs.p);
        SH_HIP(hipGetLastError());
        SH_HIP(hipMemcpyAsync(row_sums, rs.p, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    }
    SH_HIP(hipMemcpyAsync(total, tot.p, 8, hipMemcpyDeviceToHost, s));
    SH_HIP(hipStreamSynchronize(s));
    return STORM_HIP_OK;
}

}  // extern "C"
