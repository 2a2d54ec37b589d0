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
// of SW words (SW = 4, 8 or 16), every thread stages one row's packed bits in LDS
// (double buffered, one barrier per stage) exactly as loaded, and every wave
// expands the packed bits of its own fragments into 0/1 elements in registers
// right before its MFMAs. The next stage's global loads are issued before the
// current stage's MFMAs and consumed one stage later, so their latency hides
// behind a stage of matrix work.
//
// Bits are assigned to matrix K positions freely: A and B fragments place the
// same bit positions in the same K slots, so each MFMA sums products over the
// same set of positions whichever order the matrix core assigns to them. Lane
// group g (lanes 32g..32g+31) takes dwords g*SW .. g*SW+SW-1 of the stage's
// 2*SW dwords (natural memory order), 32 bits per lane group per MFMA step:
//   FP4: step 4r+d takes bit slice d of chunks 4r..4r+3: dword k, nibble n =
//        bit 4n+d of chunk 4r+k, value 2^(d-1) (d = 3 shifted onto bit 2),
//        block scales 2^(1-d) on A and B (one mask, or shift and mask, each)
//   I8 : chunk c, two K=32 steps h = 0, 1; dword d (0..3) of step h =
//        (w >> (4h+d)) & 0x01010101                   (byte j = bit 8j+4h+d)
#include "device_common.h"
#include "general.h"

namespace stormhip {

typedef int8_t i8x16 __attribute__((ext_vector_type(16)));
typedef int32_t i32x16 __attribute__((ext_vector_type(16)));
typedef int32_t i32x8 __attribute__((ext_vector_type(8)));
typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const u32x4 global_u32x4;

namespace {
constexpr int kTile = 256;        // pairs block: kTile x kTile
constexpr int kUnitScale = 127;   // E8M0 exponent of 1.0

// FP4 operand of one MFMA: bit slice d of four 32-bit chunks v.x..v.w -- dword
// k holds chunk k's bits 4n+d at nibble n, left where they are by a mask, so
// every element is the same e2m1 value 2^(d-1) (0.5, 1, 2; the sign bit d = 3
// is shifted onto 2). A and B slice alike and the MFMA's E8M0 block scales
// (2^(1-d) on each side) make every product of two set bits exactly 1.
// 5 VALU operations per four chunks and operand row.
__device__ inline i32x8 slice_fp4(const uint4& v, int d) {
    const uint32_t m = d == 0 ? 0x11111111u : d == 1 ? 0x22222222u : 0x44444444u;
    const uint32_t s = d == 3 ? 1u : 0u;
    return i32x8{(int)((v.x >> s) & m), (int)((v.y >> s) & m), (int)((v.z >> s) & m),
                 (int)((v.w >> s) & m), 0, 0, 0, 0};
}
__device__ inline i8x16 expand_i8(uint32_t w, int g) {
    const uint32_t s = 4u * g, m = 0x01010101u;
    const i32x4 d = {(int)((w >> s) & m), (int)((w >> (s + 1)) & m), (int)((w >> (s + 2)) & m),
                     (int)((w >> (s + 3)) & m)};
    return __builtin_bit_cast(i8x16, d);
}
__device__ inline uint32_t comp(const uint4& v, int c) {
    return c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w;
}

}  // namespace

template <int SW>
struct BBMfmaSmem {
    // LDS dwords per staged row: 2*SW data + 4 pad (conflict-free 16-byte
    // reads down a column of rows).
    static constexpr int kRowDw = 2 * SW + 4;
    // Packed bits of one stage: [buffer][row 0..511 (A rows, then B rows)][dw].
    uint32_t st[2][2 * kTile][kRowDw];
};

// One 256 x 256 block over stages [s0, s0 + n_stages) of SW words. With
// `accumulate`, counts are atomically added to `tri` (a K-split block: other
// blocks add the other stage ranges of the same pairs into a zeroed triangle).
// THR (thresholded queue launches): the pair counts of the block go to `tri`
// as a dense 256 x 256 row-major buffer (local row, local column); the caller
// filters them. Filtering here would hold the test's registers next to the
// accumulators.
// RECT (rectangular X . Y^T): the B rows come from `Rx->Y` (same row stride),
// every (i < n, j < Rx->ny) is a pair, and counts go to Rx->out[i*ny + j].
struct RectArgs {
    const uint64_t* Y;
    uint32_t ny;
    uint32_t* out;
};

template <bool FP4, int SW, int NW, bool THR = false, bool RECT = false>
__device__ __forceinline__ uint64_t bb_mfma_tile(BBMfmaSmem<SW>& sm, const uint64_t* X, uint32_t ld,
                                                 uint32_t s0, uint32_t n_stages,
                                                 const uint32_t* __restrict__ rowmap,
                                                 const uint32_t* __restrict__ perm, uint32_t n,
                                                 uint32_t bi, uint32_t bj,
                                                 uint32_t* __restrict__ tri, bool accumulate,
                                                 const RectArgs* __restrict__ Rx = nullptr,
                                                 bool block_local = false) {
    using Acc = typename std::conditional<FP4, f32x16, i32x16>::type;
    constexpr int L = SW / 2;        // 16-byte loads per row per stage
    constexpr int R = SW / 4;        // 16-byte LDS reads per fragment row per stage
    constexpr int NB = 16 / NW;      // 32-column MFMA blocks per wave: 2 (8 waves) or 4 (4 waves)
    constexpr int WC = kTile / 32 / NB;   // waves across the columns: 4 or 2
    constexpr int THREADS = 64 * NW;
    const uint32_t i0 = bi * kTile, j0 = bj * kTile;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wr = wave / WC, wc = wave % WC;   // 128 rows x 32*NB columns
    const int lr = lane & 31, lg = lane >> 5;

    // Thread t stages row t of the 512 staged rows (A panel, then B) and, with
    // 4 waves, row t + 256 as well.
    auto src_of = [&](uint32_t r) {
        const uint32_t pos = r < (uint32_t)kTile ? i0 + r : j0 + (r - kTile);
        const uint64_t phys = rowmap ? rowmap[pos] : (uint64_t)pos;
        const uint64_t* base = X;
        if constexpr (RECT) base = r < (uint32_t)kTile ? X : Rx->Y;
        // Rows live in global memory: loads through this pointer are global loads.
        return (global_u32x4*)(base + phys * ld) + (size_t)L * s0;
    };
    global_u32x4* src = src_of(t);
    global_u32x4* src2 = THREADS < 512 ? src_of(t + THREADS) : src;

    // Accumulators as two arrays of 4 x 2 blocks (one array of 16 blocks is
    // too large for the compiler to keep in registers; it goes to scratch).
    Acc acc_lo[4][2], acc_hi[4][NB > 2 ? NB - 2 : 1];
    auto acc = [&](int m, int q) -> Acc& { return q < 2 ? acc_lo[m][q] : acc_hi[m][q - 2]; };
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int q = 0; q < NB; ++q) acc(m, q) = Acc{};

    // The next stage's words, in named registers (an array here would be
    // placed in memory by the compiler).
    u32x4 g0, g1, g2, g3, g4, g5, g6, g7, h0, h1, h2, h3;
    auto gload = [&](uint32_t s) {
        global_u32x4* p = src + (size_t)L * s;
        g0 = p[0];
        g1 = p[1];
        if constexpr (L >= 4) { g2 = p[2]; g3 = p[3]; }
        if constexpr (L == 8) { g4 = p[4]; g5 = p[5]; g6 = p[6]; g7 = p[7]; }
        if constexpr (THREADS < 512) {
            global_u32x4* p2 = src2 + (size_t)L * s;
            h0 = p2[0];
            h1 = p2[1];
            if constexpr (L == 4) { h2 = p2[2]; h3 = p2[3]; }
        }
    };
    auto stage_store = [&](int buf) {
        u32x4* d = reinterpret_cast<u32x4*>(&sm.st[buf][t][0]);
        d[0] = g0;
        d[1] = g1;
        if constexpr (L >= 4) { d[2] = g2; d[3] = g3; }
        if constexpr (L == 8) { d[4] = g4; d[5] = g5; d[6] = g6; d[7] = g7; }
        if constexpr (THREADS < 512) {
            u32x4* d2 = reinterpret_cast<u32x4*>(&sm.st[buf][t + THREADS][0]);
            d2[0] = h0;
            d2[1] = h1;
            if constexpr (L == 4) { d2[2] = h2; d2[3] = h3; }
        }
    };

    __syncthreads();   // LDS free (previous tile of a persistent workgroup)
    gload(0);
    stage_store(0);
    if (n_stages > 1) gload(1);
    __syncthreads();
    for (uint32_t s = 0; s < n_stages; ++s) {
        const int buf = s & 1;
        // This wave's packed fragments of the stage: 4 A rows, NB B rows, each
        // SW chunks of this lane group's 32 bits. Stages of 16 words read them
        // one group of four chunks at a time (JIT), right before their MFMAs,
        // so the fragments of a whole stage never hold registers together.
        constexpr bool kJit = SW > 8;
        constexpr int RU = kJit ? 1 : R;
        uint4 pa[4][RU], pb[NB][RU];
        auto read_frags = [&](int r0) {
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int r = 0; r < RU; ++r)
                    pa[m][r] = *reinterpret_cast<const uint4*>(
                        &sm.st[buf][wr * 128 + m * 32 + lr][lg * SW + 4 * (r0 + r)]);
#pragma unroll
            for (int q = 0; q < NB; ++q)
#pragma unroll
                for (int r = 0; r < RU; ++r)
                    pb[q][r] = *reinterpret_cast<const uint4*>(
                        &sm.st[buf][kTile + wc * 32 * NB + q * 32 + lr][lg * SW + 4 * (r0 + r)]);
        };
        if constexpr (!kJit) read_frags(0);
        // Stage the next words while this stage computes: before the MFMAs
        // (4- and 8-word stages) or, with 16-word stages, right after the
        // first chunk group's MFMAs are issued, so the stage opens on matrix
        // work and the LDS stores overlap it (config 2: 150.9 -> 146 ms; after
        // the second group no gain, after the last the reload comes too late
        // and the kernel halves its speed).
        if (!kJit && s + 1 < n_stages) {
            stage_store(buf ^ 1);
            if (s + 2 < n_stages) gload(s + 2);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if constexpr (kJit) read_frags(r);
            if (kJit && r == 1 && s + 1 < n_stages) {
                stage_store(buf ^ 1);
                if (s + 2 < n_stages) gload(s + 2);
            }
            const int rr = kJit ? 0 : r;   // fragment index of chunk group r
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                if constexpr (FP4) {
                    // Step c = 4r + d: bit slice d of the four chunks 4r..4r+3.
                    const int sc = kUnitScale + 1 - (d < 2 ? d : 2);
                    i32x8 fb[NB];
#pragma unroll
                    for (int q = 0; q < NB; ++q) fb[q] = slice_fp4(pb[q][rr], d);
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        const i32x8 fa = slice_fp4(pa[m][rr], d);
#pragma unroll
                        for (int q = 0; q < NB; ++q)
                            acc(m, q) = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                                fa, fb[q], acc(m, q), 4, 4, 0, sc, 0, sc);
                    }
                } else {
                    // Chunk c = 4r + d: 32 bits per lane group = two K=32 steps
                    // (h = 0, 1 select which nibble half of each byte position
                    // the lane takes).
                    uint32_t wa[4], wb[NB];
#pragma unroll
                    for (int m = 0; m < 4; ++m) wa[m] = comp(pa[m][rr], d);
#pragma unroll
                    for (int q = 0; q < NB; ++q) wb[q] = comp(pb[q][rr], d);
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        i8x16 fb[NB];
#pragma unroll
                        for (int q = 0; q < NB; ++q) fb[q] = expand_i8(wb[q], h);
#pragma unroll
                        for (int m = 0; m < 4; ++m) {
                            const i8x16 fa = expand_i8(wa[m], h);
#pragma unroll
                            for (int q = 0; q < NB; ++q)
                                acc(m, q) = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa, fb[q], acc(m, q), 0, 0, 0);
                        }
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
            for (int q = 0; q < NB; ++q) {
                const uint32_t j = j0 + wc * 32 * NB + q * 32 + lr;
                if constexpr (RECT) {
                    if (j >= Rx->ny) continue;
                } else {
                    if (j >= n || j <= i) continue;
                }
                const uint32_t val = (uint32_t)acc(m, q)[v];
                local += val;
                if constexpr (RECT) {
                    if (Rx->out) {
                        uint32_t* dst = &Rx->out[(uint64_t)i * Rx->ny + j];
                        if (accumulate) { if (val) atomicAdd(dst, val); }
                        else            *dst = val;
                    }
                } else if (THR || block_local) {
                    tri[(i - i0) * kTile + (j - j0)] = val;
                } else if (tri) {
                    const uint32_t oj = perm ? perm[j] : j;
                    uint32_t* dst = &tri[oi < oj ? tri_index(oi, oj, n) : tri_index(oj, oi, n)];
                    if (accumulate) { if (val) atomicAdd(dst, val); }
                    else            *dst = val;
                }
            }
        }
    return local;
}

template <bool FP4, int SW, int NW, bool THR = false>
__global__ __launch_bounds__(64 * NW, 1) void bb_mfma_kernel(
    const uint64_t* __restrict__ X, uint32_t ld, uint32_t n_stages, uint32_t k_split,
    uint32_t band, const uint32_t* __restrict__ rowmap, const uint32_t* __restrict__ perm,
    uint32_t n, uint32_t T, const DevCorpus* __restrict__ Ds, const QItem* __restrict__ queue,
    const unsigned int* __restrict__ qcount, unsigned int* __restrict__ qhead,
    uint32_t* __restrict__ tri, unsigned long long* __restrict__ sum,
    unsigned long long* __restrict__ tile_ticks, uint32_t* __restrict__ partial) {
    __shared__ __attribute__((aligned(16))) BBMfmaSmem<SW> sm;
    __shared__ uint32_t s_t;
    uint64_t local = 0;
    if (THR && !qcount) return;   // thresholded instances run in queue mode only
    if (!qcount) {
        // Grid mode: position p = slice p % k_split of tile p / k_split, tiles
        // in row-major upper-triangle order (band 0) or the banded order.
        const uint32_t p = band ? xcd_position(blockIdx.x, gridDim.x) : blockIdx.x;
        uint32_t bi, bj;
        if (band) banded_tile(p / k_split, T, band, bi, bj);
        else      tri_tile(p / k_split, T, bi, bj);
        const uint32_t slice = p % k_split;
        const uint32_t per = (n_stages + k_split - 1) / k_split;
        // Clamp first so the stage count never goes through a wrapped difference.
        const uint32_t s0 = min(slice * per, n_stages);
        const uint32_t cnt = min(per, n_stages - s0);
        if (cnt && partial) {
            // K slices with a partial buffer: slice `slice` stores its block's
            // counts at (slice, tile) in the block's own layout, and a pass
            // after the launch adds the slices (plain stores, no atomics).
            const uint64_t tiles = (uint64_t)T * (T + 1) / 2;
            uint32_t* buf = partial + ((uint64_t)slice * tiles + tri_tile_id(bi, bj, T)) * kTile * kTile;
            local = bb_mfma_tile<FP4, SW, NW>(sm, X, ld, s0, cnt, rowmap, perm, n, bi, bj, buf,
                                              false, nullptr, true);
        } else if (cnt) {
            local = bb_mfma_tile<FP4, SW, NW>(sm, X, ld, s0, cnt, rowmap, perm, n, bi, bj, tri,
                                              k_split > 1);
        }
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
            const uint32_t ns = max(1u, (D.nw + SW - 1) / SW);
            uint64_t part;
            if constexpr (THR) {
                // The block's counts through its own 256 x 256 buffer, then the
                // Jaccard test per pair in a plain loop.
                uint32_t* buf = tri + (uint64_t)blockIdx.x * kTile * kTile;
                part = bb_mfma_tile<FP4, SW, NW, true>(sm, D.bitmap, D.ld, 0, ns, D.rowmap,
                                                       D.perm, D.n, e.bi, e.bj, buf, false);
                __threadfence();
                __syncthreads();
                const uint32_t i0 = e.bi * kTile, j0 = e.bj * kTile;
                for (uint32_t k = threadIdx.x; k < kTile * kTile; k += blockDim.x) {
                    const uint32_t i = i0 + k / kTile, j = j0 + k % kTile;
                    if (i >= D.n || j >= D.n || j <= i) continue;
                    emit_pair(Ds[e.inst], D.perm[i], D.perm[j], buf[k]);
                }
            } else {
                // K slice `slice` of ks (sub 0: the whole tile); slices add
                // into the triangle the caller zeroed.
                const uint32_t ks = max(1u, e.sub >> 16), slice = e.sub & 0xFFFFu;
                const uint32_t per = (ns + ks - 1) / ks;
                const uint32_t s0 = min(slice * per, ns);
                const uint32_t cnt = min(per, ns - s0);
                // A position-packed triangle takes the tile's positions as
                // they are (no map to original rows).
                const uint32_t* out_perm = D.tri_pos ? nullptr : D.perm;
                part = cnt ? bb_mfma_tile<FP4, SW, NW>(sm, D.bitmap, D.ld, s0, cnt, D.rowmap,
                                                       out_perm, D.n, e.bi, e.bj, D.tri, ks > 1)
                           : 0;
            }
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

// Rectangular X . Y^T over Tx x Ty tiles of 256 x 256 pairs, row-major, each
// tile's stages split over k_split blocks (which then add into a zeroed
// output); block b takes its XCD-aware position so each XCD sweeps a run of
// tiles sharing X panels. 8-word stages, 8 waves.
template <bool FP4>
__global__ __launch_bounds__(512, 1) void bb_mfma_rect_kernel(
    const uint64_t* __restrict__ X, uint32_t nx, const uint64_t* __restrict__ Y, uint32_t ny,
    uint32_t ld, uint32_t n_stages, uint32_t k_split, uint32_t Ty, uint32_t* __restrict__ out,
    unsigned long long* __restrict__ sum) {
    __shared__ __attribute__((aligned(16))) BBMfmaSmem<8> sm;
    const uint32_t p = xcd_position(blockIdx.x, gridDim.x);
    const uint32_t tile = p / k_split, slice = p % k_split;
    const uint32_t per = (n_stages + k_split - 1) / k_split;
    const uint32_t s0 = min(slice * per, n_stages);
    const uint32_t cnt = min(per, n_stages - s0);
    const RectArgs R{Y, ny, out};
    uint64_t local = 0;
    if (cnt)
        local = bb_mfma_tile<FP4, 8, 8, false, true>(sm, X, ld, s0, cnt, nullptr, nullptr, nx,
                                                     tile / Ty, tile % Ty, nullptr, k_split > 1,
                                                     &R);
    local = wave_sum_u64(local);
    if ((threadIdx.x & 63) == 0 && local) atomicAdd(sum, (unsigned long long)local);
}

// Exactness and layout checks of one bitmap matrix (n_words per row, rows
// ld words apart) for the matrix-core kernel with this stage width.
int bb_mfma_check(bool fp4, uint32_t sw, uint32_t n_words, uint32_t ld, uint32_t acc_words) {
    if (acc_words == 0) acc_words = n_words;
    if (sw != 4 && sw != 8 && sw != 16) {
        set_error("bb_mfma: stage width %u words (4, 8 or 16)", sw);
        return STORM_HIP_ERR_INVALID;
    }
    const uint32_t n_stages = (n_words + sw - 1) / sw;
    // Two stages are read before the loop tests its bound: rows must hold them.
    if ((uint64_t)std::max(n_stages, 2u) * sw > ld) {
        set_error("bb_mfma: %u stages of %u words exceed row stride %u", n_stages, sw, ld);
        return STORM_HIP_ERR_INVALID;
    }
    // Exactness bounds on what one item accumulates: i32 accumulators below
    // 2^31 bits; f32 below 2^24.
    if (acc_words >= (fp4 ? (1u << 18) : (1u << 25))) {
        set_error("bb_mfma: %u words per item too many for exact %s accumulation", acc_words,
                  fp4 ? "f32" : "i32");
        return STORM_HIP_ERR_RANGE;
    }
    return STORM_HIP_OK;
}

namespace {

struct GridLaunch {
    hipStream_t s;
    uint32_t g;
    const uint64_t* X;
    uint32_t ld, ns, k_split, band;
    const uint32_t *rowmap, *perm;
    uint32_t n, T;
    uint32_t* tri;
    unsigned long long* sum;
    uint32_t* partial;
    template <bool FP4, int SW, int NW>
    void operator()() const {
        hipLaunchKernelGGL((bb_mfma_kernel<FP4, SW, NW>), dim3(g), dim3(64 * NW), 0, s, X, ld, ns,
                           k_split, band, rowmap, perm, n, T, nullptr, nullptr, nullptr, nullptr,
                           tri, sum, nullptr, partial);
    }
};

struct QueueLaunch {
    hipStream_t s;
    uint32_t grid;
    const DevCorpus* Ds;
    const QItem* queue;
    const unsigned int* qcount;
    unsigned int* qhead;
    unsigned long long* tile_ticks;
    template <bool FP4, int SW, int NW>
    void operator()() const {
        hipLaunchKernelGGL((bb_mfma_kernel<FP4, SW, NW>), dim3(grid), dim3(64 * NW), 0, s, nullptr,
                           0u, 1u, 1u, 0u, nullptr, nullptr, 0u, 0u, Ds, queue, qcount, qhead,
                           nullptr, nullptr, tile_ticks, nullptr);
    }
};

struct Occupancy {
    int* nb;
    hipError_t* e;
    template <bool FP4, int SW, int NW>
    void operator()() const {
        *e = hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, bb_mfma_kernel<FP4, SW, NW>, 64 * NW, 0);
    }
};

// Shape dispatch: (format, stage words, waves) -> one kernel instance.
template <class F>
int with_shape(bool fp4, const MfmaConfig& cfg, F&& f) {
    const uint32_t sw = cfg.stage_words, nw = cfg.waves;
    if (fp4 && sw == 16 && nw == 8) { f.template operator()<true, 16, 8>(); return STORM_HIP_OK; }
    if (!fp4 && sw == 16 && nw == 8) { f.template operator()<false, 16, 8>(); return STORM_HIP_OK; }
    if (fp4 && sw == 8 && nw == 8)  { f.template operator()<true, 8, 8>(); return STORM_HIP_OK; }
    if (fp4 && sw == 4 && nw == 8)  { f.template operator()<true, 4, 8>(); return STORM_HIP_OK; }
    if (fp4 && sw == 8 && nw == 4)  { f.template operator()<true, 8, 4>(); return STORM_HIP_OK; }
    if (fp4 && sw == 4 && nw == 4)  { f.template operator()<true, 4, 4>(); return STORM_HIP_OK; }
    if (!fp4 && sw == 8 && nw == 8) { f.template operator()<false, 8, 8>(); return STORM_HIP_OK; }
    if (!fp4 && sw == 4 && nw == 8) { f.template operator()<false, 4, 8>(); return STORM_HIP_OK; }
    if (!fp4 && sw == 8 && nw == 4) { f.template operator()<false, 8, 4>(); return STORM_HIP_OK; }
    if (!fp4 && sw == 4 && nw == 4) { f.template operator()<false, 4, 4>(); return STORM_HIP_OK; }
    set_error("bb_mfma: no kernel for %u-word stages and %u waves", sw, nw);
    return STORM_HIP_ERR_INVALID;
}

// Zero fill of a u32 array as a kernel.
__global__ void zero_u32_kernel(uint32_t* p, uint64_t count) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
         k += (uint64_t)gridDim.x * blockDim.x)
        p[k] = 0;
}

}  // namespace

// k_split > 1 (grid mode only): each tile's stages are split over k_split
// blocks that atomically add into `tri`, which the caller has zeroed.
int launch_bb_mfma_fmt(bool fp4, const MfmaConfig& cfg, hipStream_t s, const uint64_t* X,
                       uint32_t ld, uint32_t n_words, const uint32_t* rowmap, const uint32_t* perm,
                       uint32_t n, uint32_t* tri, unsigned long long* sum, uint32_t k_split,
                       uint32_t* partial) {
    const uint32_t sw = cfg.stage_words;
    SH_TRY(bb_mfma_check(fp4, sw, n_words, ld));
    if (k_split == 0) k_split = 1;
    const uint32_t n_stages = (n_words + sw - 1) / sw;
    const uint32_t T = (n + kTile - 1) / kTile;
    const uint64_t g = (uint64_t)T * (T + 1) / 2 * k_split;
    if (g == 0) return STORM_HIP_OK;
    if (g > 0xFFFFFFFFull) {
        set_error("bb_mfma: %llu blocks exceed one launch", (unsigned long long)g);
        return STORM_HIP_ERR_RANGE;
    }
    const uint32_t ns = n_stages ? n_stages : 1;
    const GridLaunch L{s, (uint32_t)g, X, ld, ns, k_split, cfg.band, rowmap, perm, n, T, tri, sum,
                       k_split > 1 ? partial : nullptr};
    SH_TRY(with_shape(fp4, cfg, L));
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

uint32_t bb_valu_blocks_per_cu(int bm);

uint32_t bb_blocks_per_cu(uint32_t tile, const MfmaConfig& cfg) {
    if (tile != 256) return bb_valu_blocks_per_cu((int)tile);
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
    if (with_shape(true, cfg, Occupancy{&nb, &e}) != STORM_HIP_OK) return 1;
    return (e == hipSuccess && nb > 0) ? (uint32_t)nb : 1;
}

// Queue mode over the corpora in Ds. The caller has checked every corpus that
// can queue a tile with bb_mfma_check (and picks fp4 only when all allow it).
// Thresholded calls (thr_buf: grid x 256 x 256 u32 of block buffers) run the
// 8-word, 8-wave shape's thresholded instance.
int launch_bb_mfma_queue(bool fp4, const MfmaConfig& cfg, hipStream_t s, const DevCorpus* Ds,
                         const QItem* queue, const unsigned int* qcount, unsigned int* qhead,
                         uint32_t grid, unsigned long long* tile_ticks, uint32_t* thr_buf) {
    if (grid == 0) return STORM_HIP_OK;
    if (thr_buf) {
        if (fp4)
            hipLaunchKernelGGL((bb_mfma_kernel<true, 8, 8, true>), dim3(grid), dim3(512), 0, s,
                               nullptr, 0u, 1u, 1u, 0u, nullptr, nullptr, 0u, 0u, Ds, queue,
                               qcount, qhead, thr_buf, nullptr, tile_ticks, nullptr);
        else
            hipLaunchKernelGGL((bb_mfma_kernel<false, 8, 8, true>), dim3(grid), dim3(512), 0, s,
                               nullptr, 0u, 1u, 1u, 0u, nullptr, nullptr, 0u, 0u, Ds, queue,
                               qcount, qhead, thr_buf, nullptr, tile_ticks, nullptr);
    } else {
        SH_TRY(with_shape(fp4, cfg, QueueLaunch{s, grid, Ds, queue, qcount, qhead, tile_ticks}));
    }
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

// Rectangular launch: rows of X (nx, padded to 256) against rows of Y (ny,
// padded to 256), same row stride; out (nx * ny, row-major) may be null.
int launch_bb_rect(bool fp4, const MfmaConfig& cfg, hipStream_t s, const uint64_t* X, uint32_t nx,
                   const uint64_t* Y, uint32_t ny, uint32_t ld, uint32_t n_words, uint32_t* out,
                   unsigned long long* sum, uint32_t cu_count) {
    SH_TRY(bb_mfma_check(fp4, 8, n_words, ld));
    const uint64_t Tx = (nx + kTile - 1) / kTile, Ty = (ny + kTile - 1) / kTile;
    const uint64_t tiles = Tx * Ty;
    if (tiles == 0) return STORM_HIP_OK;
    const uint32_t n_stages = std::max(1u, (n_words + 7) / 8);
    const uint64_t want = (uint64_t)std::max(1u, cfg.blocks_per_cu) * cu_count;
    uint32_t k_split = 1;
    if (tiles < want) {
        k_split = (uint32_t)std::min<uint64_t>((want + tiles - 1) / tiles,
                                               std::max(1u, n_stages / 4));
        k_split = std::max(1u, k_split);
        const uint32_t per = (n_stages + k_split - 1) / k_split;
        k_split = (n_stages + per - 1) / per;
    }
    const uint64_t g = tiles * k_split;
    if (g > 0xFFFFFFFFull) {
        set_error("bb_mfma: %llu blocks exceed one launch", (unsigned long long)g);
        return STORM_HIP_ERR_RANGE;
    }
    if (k_split > 1 && out) {
        const uint64_t count = (uint64_t)nx * ny;
        hipLaunchKernelGGL(zero_u32_kernel, dim3((uint32_t)std::min<uint64_t>((count + 255) / 256, 65536)),
                           dim3(256), 0, s, out, count);
        SH_HIP(hipGetLastError());
    }
    if (fp4)
        hipLaunchKernelGGL(bb_mfma_rect_kernel<true>, dim3((uint32_t)g), dim3(512), 0, s, X, nx, Y,
                           ny, ld, n_stages, k_split, (uint32_t)Ty, out, sum);
    else
        hipLaunchKernelGGL(bb_mfma_rect_kernel<false>, dim3((uint32_t)g), dim3(512), 0, s, X, nx, Y,
                           ny, ld, n_stages, k_split, (uint32_t)Ty, out, sum);
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

// Sums the K slices' block-local counts of every pair into the triangle: a
// workgroup per (tile, tile row), a thread per column; slices are read
// 256 columns at a time and each row segment of the triangle is written once.
__global__ __launch_bounds__(256) void ksplit_sum_kernel(const uint32_t* __restrict__ partial,
                                                         uint32_t k_split, uint32_t T, uint32_t n,
                                                         const uint32_t* __restrict__ perm,
                                                         uint32_t* __restrict__ tri) {
    const uint64_t tiles = (uint64_t)T * (T + 1) / 2;
    const uint64_t t = blockIdx.x / kTile;
    const uint32_t r = blockIdx.x % kTile, c = threadIdx.x;
    uint32_t bi, bj;
    tri_tile(t, T, bi, bj);
    const uint32_t i = bi * kTile + r, j = bj * kTile + c;
    if (i >= n || j >= n || j <= i) return;
    const uint32_t* src = partial + t * kTile * kTile + (uint64_t)r * kTile + c;
    uint32_t v = 0;
    for (uint32_t sl = 0; sl < k_split; ++sl) v += src[(uint64_t)sl * tiles * kTile * kTile];
    const uint32_t oi = perm ? perm[i] : i, oj = perm ? perm[j] : j;
    tri[oi < oj ? tri_index(oi, oj, n) : tri_index(oj, oi, n)] = v;
}

int launch_bb_dense(uint32_t bb_kernel, const MfmaConfig& cfg, hipStream_t s, const uint64_t* X,
                    uint32_t ld, uint32_t n_words, const uint32_t* rowmap, uint32_t n,
                    uint32_t* tri, unsigned long long* sum, uint32_t cu_count,
                    const uint32_t* perm, bool may_allocate) {
    if (bb_kernel == STORM_HIP_BB_VALU)
        return launch_bb_valu(s, 128, X, ld, n_words, rowmap, perm, n, tri, sum);
    const bool fp4 = bb_kernel == STORM_HIP_BB_MFMA_FP4 ||
                     (bb_kernel == STORM_HIP_BB_AUTO && n_words < (1u << 18));
    // A caller's row stride shorter than two stages of the configured width
    // (rows of a few words, laid out tightly) takes the widest stage it holds.
    MfmaConfig shape = cfg;
    auto fits = [&](uint32_t sw) {
        return (uint64_t)std::max((n_words + sw - 1) / sw, 2u) * sw <= ld;
    };
    while (shape.stage_words > 4 && !fits(shape.stage_words)) shape.stage_words /= 2;
    // Too few 256 x 256 blocks to fill the chip: split K so that the blocks
    // fit in one round of the workgroup slots (blocks_per_cu per CU, at most
    // what the shape's LDS lets reside), each stage range at least 4 stages
    // long. Rounding the split up instead leaves a second round of a few
    // blocks that costs as much as the first.
    const uint32_t T = (n + kTile - 1) / kTile;
    const uint64_t tiles = (uint64_t)T * (T + 1) / 2;
    const uint32_t n_stages = (n_words + shape.stage_words - 1) / shape.stage_words;
    const uint32_t resident = bb_blocks_per_cu(kTile, shape);
    const uint64_t want = (uint64_t)std::max(1u, std::min(shape.blocks_per_cu, resident)) * cu_count;
    uint32_t k_split = 1;
    if (tiles && n_stages && tiles < want) {
        k_split = (uint32_t)std::min<uint64_t>(want / tiles, std::max(1u, n_stages / 4));
        k_split = std::max(1u, k_split);
        // No empty slices: as many slices as ceil(n_stages / per) needs.
        const uint32_t per = (n_stages + k_split - 1) / k_split;
        k_split = (n_stages + per - 1) / per;
    }
    if (k_split > 1 && tri && may_allocate) {
        // The slices store their blocks (k_split x tiles x 256 KB, at most a
        // few hundred MB: only small corpora split K) and one pass adds them:
        // no atomic per pair per slice into a zeroed triangle.
        uint32_t* partial = nullptr;
        const size_t bytes = (size_t)k_split * tiles * kTile * kTile * 4;
        SH_HIP(hipMallocAsync((void**)&partial, bytes, s));
        int rc = launch_bb_mfma_fmt(fp4, shape, s, X, ld, n_words, rowmap, perm, n, tri, sum,
                                    k_split, partial);
        if (rc == STORM_HIP_OK) {
            hipLaunchKernelGGL(ksplit_sum_kernel, dim3((uint32_t)(tiles * kTile)), dim3(kTile), 0,
                               s, partial, k_split, T, n, perm, tri);
            if (hipGetLastError() != hipSuccess) {
                set_error("ksplit_sum_kernel launch failed");
                rc = STORM_HIP_ERR_DEVICE;
            }
        }
        SH_HIP(hipFreeAsync(partial, s));
        return rc;
    }
    if (k_split > 1 && tri) {
        const uint64_t count = (uint64_t)n * (n - 1) / 2;
        hipLaunchKernelGGL(zero_u32_kernel, dim3((uint32_t)std::min<uint64_t>((count + 255) / 256, 65536)),
                           dim3(256), 0, s, tri, count);
        SH_HIP(hipGetLastError());
    }
    return launch_bb_mfma_fmt(fp4, shape, s, X, ld, n_words, rowmap, perm, n, tri, sum, k_split);
}

}  // namespace stormhip
