// B x B cell, matrix-core form: exact |Xi n Xj| as a 0/1 Gram matrix on MFMA.
//
// Each bit of a row becomes one matrix element equal to 0 or 1, so
// |Xi n Xj| = sum_k Xi[k]*Xj[k]. Two element formats:
//   I8   int8 elements, v_mfma_i32_32x32x32_i8, i32 accumulation: exact for any
//        count below 2^31.
//   FP4  e2m1 elements (1.0 = 0b0010), v_mfma_scale_f32_32x32x64_f8f6f4 with
//        unit scales, f32 accumulation: every partial sum is an integer no
//        larger than the pair's final count, so the result is exact while
//        counts stay <= 2^24 (the launcher refuses universes above that).
//        Twice the K per instruction of I8 at the same issue rate.
//
// Structure: a workgroup of 8 waves owns a 256 x 256 block of pairs; wave w
// owns rows 128*(w/4) .. +128 and columns 64*(w%4) .. +64 of it (4 x 2 MFMA
// blocks of 32 x 32, 128 accumulator registers, two waves per SIMD). Per stage
// of 4 words, every thread stages one row's packed bits in LDS (double
// buffered, one barrier per stage), and every wave expands the packed bits of
// its own fragments into 0/1 elements in registers right before its MFMAs.
//
// The expansion is cheap because the element ORDER inside a fragment is free:
// A and B fragments are built by the same function from the same bit
// positions, so each MFMA sums products over the same set of positions
// whichever order the matrix core assigns to them. For a 32-bit chunk w:
//   FP4: dword d (0..3) = (w >> (d-1)) & 0x22222222   (nibble n = bit 4n+d)
//   I8 : two K=32 steps h = 0, 1 per chunk; dword d (0..3) of step h =
//        (w >> (4h+d)) & 0x01010101                   (byte j = bit 8j+4h+d)
// i.e. one shift and one AND per output dword.
#include "device_common.h"
#include "general.h"

namespace stormhip {

typedef int8_t i8x16 __attribute__((ext_vector_type(16)));
typedef int32_t i32x16 __attribute__((ext_vector_type(16)));
typedef int32_t i32x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace {
constexpr int kTile = 256;        // pairs block: kTile x kTile
constexpr int kThreads = 512;     // 8 waves
constexpr int kWordsStage = 4;    // packed 64-bit words per stage per row
constexpr int kRowDw = 12;        // LDS dwords per staged row (8 data + 4 pad)
constexpr int kUnitScale = 127;   // E8M0 exponent of 1.0

__device__ inline i32x8 expand_fp4(uint32_t w) {
    return i32x8{(int)((w << 1) & 0x22222222u), (int)(w & 0x22222222u),
                 (int)((w >> 1) & 0x22222222u), (int)((w >> 2) & 0x22222222u), 0, 0, 0, 0};
}
__device__ inline i8x16 expand_i8(uint32_t w, int g) {
    const uint32_t s = 4u * g;
    const uint32_t d0 = (w >> s) & 0x01010101u, d1 = (w >> (s + 1)) & 0x01010101u;
    const uint32_t d2 = (w >> (s + 2)) & 0x01010101u, d3 = (w >> (s + 3)) & 0x01010101u;
    i8x16 r;
    uint32_t tmp[4] = {d0, d1, d2, d3};
    __builtin_memcpy(&r, tmp, 16);
    return r;
}
}  // namespace

struct BBMfmaSmem {
    // Packed bits of one stage: [buffer][row 0..511 (A rows, then B rows)][dw].
    // Row layout: dwords 0..3 = low halves of words 0..3, 4..7 = high halves.
    uint32_t st[2][kThreads][kRowDw];
};

// One 256 x 256 block over stages [s0, s0 + n_stages) of the rows. With
// `accumulate`, counts are atomically added to `tri` (a K-split block: other
// blocks add the other stage ranges of the same pairs into a zeroed triangle).
template <bool FP4>
__device__ __forceinline__ uint64_t bb_mfma_tile(BBMfmaSmem& sm, const uint64_t* __restrict__ X, uint32_t ld,
                                 uint32_t s0, uint32_t n_stages, const uint32_t* __restrict__ rowmap,
                                 const uint32_t* __restrict__ perm, uint32_t n, uint32_t bi,
                                 uint32_t bj, uint32_t* __restrict__ tri, bool accumulate) {
    using Acc = typename std::conditional<FP4, f32x16, i32x16>::type;
    const uint32_t i0 = bi * kTile, j0 = bj * kTile;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wr = wave >> 2, wc = wave & 3;   // 128 rows x 64 columns
    const int lr = lane & 31, lg = lane >> 5;

    // Thread t stages row t of the A panel (t < 256) or row t-256 of B.
    const uint32_t pos = t < kTile ? i0 + t : j0 + (t - kTile);
    const uint64_t phys = rowmap ? rowmap[pos] : (uint64_t)pos;
    const uint4* src = reinterpret_cast<const uint4*>(X + phys * ld) + 2 * (size_t)s0;

    Acc acc[4][2];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[m][q][v] = 0;

    auto stage_store = [&](int buf, const uint4& a, const uint4& b) {
        // words: a = (w0.lo, w0.hi, w1.lo, w1.hi), b = (w2.lo, w2.hi, w3.lo, w3.hi)
        *reinterpret_cast<uint4*>(&sm.st[buf][t][0]) = make_uint4(a.x, a.z, b.x, b.z);
        *reinterpret_cast<uint4*>(&sm.st[buf][t][4]) = make_uint4(a.y, a.w, b.y, b.w);
    };

    __syncthreads();   // LDS free (previous tile of a persistent workgroup)
    uint4 ga = src[0], gb = src[1];
    stage_store(0, ga, gb);
    if (n_stages > 1) { ga = src[2]; gb = src[3]; }
    __syncthreads();
    for (uint32_t s = 0; s < n_stages; ++s) {
        const int buf = s & 1;
        // This wave's packed fragments of the stage: 4 A rows, 2 B rows, each
        // 4 chunks of this lane group's 32 bits.
        uint4 pa[4], pb[2];
#pragma unroll
        for (int m = 0; m < 4; ++m)
            pa[m] = *reinterpret_cast<const uint4*>(&sm.st[buf][wr * 128 + m * 32 + lr][lg * 4]);
#pragma unroll
        for (int q = 0; q < 2; ++q)
            pb[q] = *reinterpret_cast<const uint4*>(&sm.st[buf][kTile + wc * 64 + q * 32 + lr][lg * 4]);
        // Stage the next words while this stage computes.
        if (s + 1 < n_stages) {
            stage_store(buf ^ 1, ga, gb);
            if (s + 2 < n_stages) { ga = src[2 * (s + 2)]; gb = src[2 * (s + 2) + 1]; }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            uint32_t wa[4], wb[2];
#pragma unroll
            for (int m = 0; m < 4; ++m) wa[m] = c == 0 ? pa[m].x : c == 1 ? pa[m].y : c == 2 ? pa[m].z : pa[m].w;
#pragma unroll
            for (int q = 0; q < 2; ++q) wb[q] = c == 0 ? pb[q].x : c == 1 ? pb[q].y : c == 2 ? pb[q].z : pb[q].w;
            if constexpr (FP4) {
                i32x8 fb[2];
#pragma unroll
                for (int q = 0; q < 2; ++q) fb[q] = expand_fp4(wb[q]);
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const i32x8 fa = expand_fp4(wa[m]);
#pragma unroll
                    for (int q = 0; q < 2; ++q)
                        acc[m][q] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                            fa, fb[q], acc[m][q], 4, 4, 0, kUnitScale, 0, kUnitScale);
                }
            } else {
                // 32 bits per lane group per chunk = two K=32 steps (g = 0, 1
                // select which nibble half of each byte position the lane takes).
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    i8x16 fb[2];
#pragma unroll
                    for (int q = 0; q < 2; ++q) fb[q] = expand_i8(wb[q], h);
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        const i8x16 fa = expand_i8(wa[m], h);
#pragma unroll
                        for (int q = 0; q < 2; ++q)
                            acc[m][q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa, fb[q], acc[m][q], 0, 0, 0);
                    }
                }
            }
        }
        __syncthreads();   // next buffer fully staged; this buffer fully read
    }

    // C layout of a 32x32 block: lane l, register v -> row 8*(v/4) + 4*(l/32)
    // + (v%4), column l % 32.
    uint64_t local = 0;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const uint32_t i = i0 + wr * 128 + m * 32 + 8 * (v >> 2) + 4 * lg + (v & 3);
            if (i >= n) continue;
            const uint32_t oi = perm ? perm[i] : i;
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const uint32_t j = j0 + wc * 64 + q * 32 + lr;
                if (j >= n || j <= i) continue;
                const uint32_t val = (uint32_t)acc[m][q][v];
                local += val;
                if (tri) {
                    const uint32_t oj = perm ? perm[j] : j;
                    uint32_t* dst = &tri[oi < oj ? tri_index(oi, oj, n) : tri_index(oj, oi, n)];
                    if (accumulate) { if (val) atomicAdd(dst, val); }
                    else            *dst = val;
                }
            }
        }
    return local;
}

template <bool FP4>
__global__ __launch_bounds__(kThreads, 1) void bb_mfma_kernel(
    const uint64_t* __restrict__ X, uint32_t ld, uint32_t n_stages, uint32_t k_split,
    const uint32_t* __restrict__ rowmap, const uint32_t* __restrict__ perm, uint32_t n,
    uint32_t T, const DevCorpus* __restrict__ Ds, const QItem* __restrict__ queue,
    const unsigned int* __restrict__ qcount, unsigned int* __restrict__ qhead,
    uint32_t* __restrict__ tri, unsigned long long* __restrict__ sum,
    unsigned long long* __restrict__ tile_ticks) {
    __shared__ __attribute__((aligned(16))) BBMfmaSmem sm;
    __shared__ uint32_t s_t;
    uint64_t local = 0;
    if (!qcount) {
        // Grid mode: block b is slice b % k_split of upper-triangle tile b / k_split.
        uint32_t bi, bj;
        tri_tile(blockIdx.x / k_split, T, bi, bj);
        const uint32_t slice = blockIdx.x % k_split;
        const uint32_t per = (n_stages + k_split - 1) / k_split;
        // Clamp first so the stage count never goes through a wrapped difference.
        const uint32_t s0 = min(slice * per, n_stages);
        const uint32_t cnt = min(per, n_stages - s0);
        if (cnt)
            local = bb_mfma_tile<FP4>(sm, X, ld, s0, cnt, rowmap, perm, n, bi, bj, tri, k_split > 1);
    } else {
        for (;;) {
            __syncthreads();
            if (threadIdx.x == 0) s_t = atomicAdd(qhead, 1u);
            __syncthreads();
            const uint32_t t = s_t;
            if (t >= *qcount) break;
            const QItem e = queue[t];
            const DevCorpus D = Ds[e.inst];
            const uint64_t t0 = wall_clock64();
            const uint32_t ns = max(1u, (D.nw + kWordsStage - 1) / kWordsStage);
            uint64_t part = bb_mfma_tile<FP4>(sm, D.bitmap, D.ld, 0, ns, D.rowmap, D.perm, D.n,
                                              e.bi, e.bj, D.tri, false);
            part = wave_sum_u64(part);
            if ((threadIdx.x & 63) == 0 && part) atomicAdd(D.sum, (unsigned long long)part);
            __syncthreads();
            if (threadIdx.x == 0) {
                const uint64_t now = wall_clock64();
                if (tile_ticks) tile_ticks[t] = now - t0;
                if (D.finish) atomicMax(D.finish, (unsigned long long)now);
            }
        }
        return;
    }
    local = wave_sum_u64(local);
    if ((threadIdx.x & 63) == 0 && local) atomicAdd(sum, (unsigned long long)local);
}

// k_split > 1 (grid mode only): each tile's stages are split over k_split
// blocks that atomically add into `tri`, which the caller has zeroed.
// Exactness and layout checks shared by both launch modes.
static int mfma_checks(bool fp4, uint32_t n_words, uint32_t ld) {
    const uint32_t n_stages = (n_words + kWordsStage - 1) / kWordsStage;
    // Two stages are read before the loop tests its bound: rows must hold them.
    if ((uint64_t)std::max(n_stages, 2u) * kWordsStage > ld) {
        set_error("bb_mfma: %u stages exceed row stride %u", n_stages, ld);
        return STORM_HIP_ERR_INVALID;
    }
    // Exactness bounds: i32 accumulators below 2^31 bits; f32 below 2^24.
    if (n_words >= (fp4 ? (1u << 18) : (1u << 25))) {
        set_error("bb_mfma: universe of %u words too large for exact %s accumulation", n_words,
                  fp4 ? "f32" : "i32");
        return STORM_HIP_ERR_RANGE;
    }
    return STORM_HIP_OK;
}

int launch_bb_mfma_fmt(bool fp4, hipStream_t s, const uint64_t* X, uint32_t ld, uint32_t n_words,
                       const uint32_t* rowmap, const uint32_t* perm, uint32_t n, uint32_t* tri,
                       unsigned long long* sum, uint32_t k_split) {
    SH_TRY(mfma_checks(fp4, n_words, ld));
    if (k_split == 0) k_split = 1;
    const uint32_t n_stages = (n_words + kWordsStage - 1) / kWordsStage;
    const uint32_t T = (n + kTile - 1) / kTile;
    const uint64_t g = (uint64_t)T * (T + 1) / 2 * k_split;
    if (g == 0) return STORM_HIP_OK;
    const uint32_t ns = n_stages ? n_stages : 1;
    if (fp4)
        hipLaunchKernelGGL(bb_mfma_kernel<true>, dim3((uint32_t)g), dim3(kThreads), 0, s, X, ld, ns,
                           k_split, rowmap, perm, n, T, nullptr, nullptr, nullptr, nullptr, tri,
                           sum, nullptr);
    else
        hipLaunchKernelGGL(bb_mfma_kernel<false>, dim3((uint32_t)g), dim3(kThreads), 0, s, X, ld, ns,
                           k_split, rowmap, perm, n, T, nullptr, nullptr, nullptr, nullptr, tri,
                           sum, nullptr);
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

// Queue mode over the corpora in Ds; `max_words`/`min_ld` bound every corpus
// of the call (the caller picks fp4 only when all of them allow it).
int launch_bb_mfma_queue(bool fp4, hipStream_t s, const DevCorpus* Ds, uint32_t max_words,
                         uint32_t min_ld, const QItem* queue, const unsigned int* qcount,
                         unsigned int* qhead, uint32_t grid, unsigned long long* tile_ticks) {
    SH_TRY(mfma_checks(fp4, max_words, min_ld));
    if (grid == 0) return STORM_HIP_OK;
    if (fp4)
        hipLaunchKernelGGL(bb_mfma_kernel<true>, dim3(grid), dim3(kThreads), 0, s, nullptr, 0u, 1u,
                           1u, nullptr, nullptr, 0u, 0u, Ds, queue, qcount, qhead, nullptr, nullptr,
                           tile_ticks);
    else
        hipLaunchKernelGGL(bb_mfma_kernel<false>, dim3(grid), dim3(kThreads), 0, s, nullptr, 0u, 1u,
                           1u, nullptr, nullptr, 0u, 0u, Ds, queue, qcount, qhead, nullptr, nullptr,
                           tile_ticks);
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

int launch_bb_dense(uint32_t bb_kernel, hipStream_t s, const uint64_t* X, uint32_t ld,
                    uint32_t n_words, const uint32_t* rowmap, uint32_t n, uint32_t* tri,
                    unsigned long long* sum, uint32_t cu_count) {
    if (bb_kernel == STORM_HIP_BB_VALU)
        return launch_bb_valu(s, 128, X, ld, n_words, rowmap, nullptr, n, tri, sum);
    const bool fp4 = bb_kernel == STORM_HIP_BB_MFMA_FP4 ||
                     (bb_kernel == STORM_HIP_BB_AUTO && n_words < (1u << 18));
    // Too few 256 x 256 blocks to fill the chip: split K so that about two
    // blocks per CU run, each stage range at least 8 stages long.
    const uint32_t T = (n + kTile - 1) / kTile;
    const uint64_t tiles = (uint64_t)T * (T + 1) / 2;
    const uint32_t n_stages = (n_words + kWordsStage - 1) / kWordsStage;
    uint32_t k_split = 1;
    if (tiles && n_stages && tiles < 2ull * cu_count) {
        k_split = (uint32_t)std::min<uint64_t>((2ull * cu_count + tiles - 1) / tiles,
                                               std::max(1u, n_stages / 8));
        k_split = std::max(1u, k_split);
        // No empty slices: as many slices as ceil(n_stages / per) needs.
        const uint32_t per = (n_stages + k_split - 1) / k_split;
        k_split = (n_stages + per - 1) / per;
    }
    if (k_split > 1 && tri)
        SH_HIP(hipMemsetAsync(tri, 0, (uint64_t)n * (n - 1) / 2 * 4, s));
    return launch_bb_mfma_fmt(fp4, s, X, ld, n_words, rowmap, nullptr, n, tri, sum, k_split);
}

}  // namespace stormhip
