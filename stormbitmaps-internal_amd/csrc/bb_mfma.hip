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
//        Twice the K per instruction at the same issue rate.
// The bits -> elements expansion is done ONCE per workgroup stage into LDS and
// shared by the four waves.
//
// Same contract as the VALU form (bb_valu.hip): a 256 x 256 block of pairs per
// workgroup, upper triangle of blocks (or a device queue of blocks), optional
// rowmap/perm indirection.
//
// Operand mapping: lane l of an A (or B) fragment holds 16 bytes of row
// (l % 32) of its 32-row block, taken from the byte range [16*(l/32), +16) of
// that row's expanded k-chunk. A and B use the identical mapping, so each MFMA
// sums products over the same set of bit positions whatever order the matrix
// core assigns to them.
#include "device_common.h"
#include "internal.h"

namespace stormhip {

typedef int8_t i8x16 __attribute__((ext_vector_type(16)));
typedef int32_t i32x16 __attribute__((ext_vector_type(16)));
typedef int32_t i32x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace {
constexpr int kTile = 256;       // pairs block: kTile x kTile
constexpr int kWordsStage = 2;   // packed 64-bit words per LDS stage
constexpr int kRowBytes = 32;    // expanded bytes per row per k-chunk

// I8: 4 bits -> 4 bytes of 0/1. nibble * 0x00204081 puts bit b at bit 8b
// with no two partial products overlapping; the mask keeps those bits.
__device__ inline uint32_t nib_to_bytes(uint32_t x) { return (x * 0x00204081u) & 0x01010101u; }

// I8: one 32-bit chunk -> 32 bytes (byte k = bit k) as 8 dwords.
__device__ inline void expand_i8(uint32_t d, uint4& lo, uint4& hi) {
    lo.x = nib_to_bytes(d & 0xF);
    lo.y = nib_to_bytes((d >> 4) & 0xF);
    lo.z = nib_to_bytes((d >> 8) & 0xF);
    lo.w = nib_to_bytes((d >> 12) & 0xF);
    hi.x = nib_to_bytes((d >> 16) & 0xF);
    hi.y = nib_to_bytes((d >> 20) & 0xF);
    hi.z = nib_to_bytes((d >> 24) & 0xF);
    hi.w = nib_to_bytes(d >> 28);
}

// FP4: 8 bits -> 8 e2m1 nibbles (bit n -> nibble n = 0b0010 or 0).
__device__ inline uint32_t byte_to_fp4(uint32_t x) {
    x = (x | (x << 12)) & 0x000F000Fu;
    x = (x | (x << 6)) & 0x03030303u;
    x = (x | (x << 3)) & 0x11111111u;
    return x << 1;
}

// FP4: one 64-bit chunk (lo, hi dwords) -> 32 bytes as 8 dwords.
__device__ inline void expand_fp4(uint32_t lo32, uint32_t hi32, uint4& lo, uint4& hi) {
    lo.x = byte_to_fp4(lo32 & 0xFF);
    lo.y = byte_to_fp4((lo32 >> 8) & 0xFF);
    lo.z = byte_to_fp4((lo32 >> 16) & 0xFF);
    lo.w = byte_to_fp4(lo32 >> 24);
    hi.x = byte_to_fp4(hi32 & 0xFF);
    hi.y = byte_to_fp4((hi32 >> 8) & 0xFF);
    hi.z = byte_to_fp4((hi32 >> 16) & 0xFF);
    hi.w = byte_to_fp4(hi32 >> 24);
}

constexpr int kUnitScale = 127;   // E8M0 exponent of 1.0
}  // namespace

template <bool FP4>
struct BBMfmaSmem {
    static constexpr int kChunks = FP4 ? kWordsStage : kWordsStage * 2;
    // Expanded operands of one stage: [A|B][chunk][row][32 bytes].
    uint8_t lds[2][kChunks][kTile][kRowBytes];
};

template <bool FP4>
__device__ uint64_t bb_mfma_tile(BBMfmaSmem<FP4>& sm, const uint64_t* __restrict__ X, uint32_t ld,
                                 uint32_t n_stages, const uint32_t* __restrict__ rowmap,
                                 const uint32_t* __restrict__ perm, uint32_t n, uint32_t bi,
                                 uint32_t bj, uint32_t* __restrict__ tri) {
    constexpr int kChunks = BBMfmaSmem<FP4>::kChunks;
    using Acc = typename std::conditional<FP4, f32x16, i32x16>::type;
    const uint32_t i0 = bi * kTile, j0 = bj * kTile;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wr = wave >> 1, wc = wave & 1;   // wave's 128 x 128 sub-block
    const int lr = lane & 31, lg = lane >> 5;

    // Thread t stages row t of the A panel and row t of the B panel.
    const uint64_t ra = rowmap ? rowmap[i0 + t] : (uint64_t)(i0 + t);
    const uint64_t rb = rowmap ? rowmap[j0 + t] : (uint64_t)(j0 + t);
    const uint4* pa = reinterpret_cast<const uint4*>(X + ra * ld);
    const uint4* pb = reinterpret_cast<const uint4*>(X + rb * ld);

    Acc acc[4][4];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[m][q][v] = 0;

    __syncthreads();   // LDS free (previous tile of a persistent workgroup)
    uint4 ga = pa[0], gb = pb[0];
    for (uint32_t s = 0; s < n_stages; ++s) {
        __syncthreads();   // previous stage's fragments fully read
        {
            const uint32_t da[4] = {ga.x, ga.y, ga.z, ga.w};
            const uint32_t db[4] = {gb.x, gb.y, gb.z, gb.w};
#pragma unroll
            for (int c = 0; c < kChunks; ++c) {
                uint4 lo, hi;
                if constexpr (FP4) expand_fp4(da[2 * c], da[2 * c + 1], lo, hi);
                else               expand_i8(da[c], lo, hi);
                *reinterpret_cast<uint4*>(&sm.lds[0][c][t][0]) = lo;
                *reinterpret_cast<uint4*>(&sm.lds[0][c][t][16]) = hi;
                if constexpr (FP4) expand_fp4(db[2 * c], db[2 * c + 1], lo, hi);
                else               expand_i8(db[c], lo, hi);
                *reinterpret_cast<uint4*>(&sm.lds[1][c][t][0]) = lo;
                *reinterpret_cast<uint4*>(&sm.lds[1][c][t][16]) = hi;
            }
        }
        __syncthreads();
        if (s + 1 < n_stages) { ga = pa[s + 1]; gb = pb[s + 1]; }
#pragma unroll
        for (int c = 0; c < kChunks; ++c) {
            i32x8 fa[4], fb[4];   // I8 uses the low 4 dwords only
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const uint4 v = *reinterpret_cast<const uint4*>(&sm.lds[0][c][wr * 128 + m * 32 + lr][lg * 16]);
                fa[m] = i32x8{(int)v.x, (int)v.y, (int)v.z, (int)v.w, 0, 0, 0, 0};
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint4 v = *reinterpret_cast<const uint4*>(&sm.lds[1][c][wc * 128 + q * 32 + lr][lg * 16]);
                fb[q] = i32x8{(int)v.x, (int)v.y, (int)v.z, (int)v.w, 0, 0, 0, 0};
            }
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if constexpr (FP4) {
                        acc[m][q] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                            fa[m], fb[q], acc[m][q], 4, 4, 0, kUnitScale, 0, kUnitScale);
                    } else {
                        const i32x16 cur = acc[m][q];
                        i8x16 a8, b8;
                        __builtin_memcpy(&a8, &fa[m], 16);
                        __builtin_memcpy(&b8, &fb[q], 16);
                        acc[m][q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a8, b8, cur, 0, 0, 0);
                    }
                }
        }
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
            for (int q = 0; q < 4; ++q) {
                const uint32_t j = j0 + wc * 128 + q * 32 + lr;
                if (j >= n || j <= i) continue;
                const uint32_t val = (uint32_t)acc[m][q][v];
                local += val;
                if (tri) {
                    const uint32_t oj = perm ? perm[j] : j;
                    tri[oi < oj ? tri_index(oi, oj, n) : tri_index(oj, oi, n)] = val;
                }
            }
        }
    return local;
}

template <bool FP4>
__global__ __launch_bounds__(256, 1) void bb_mfma_kernel(
    const uint64_t* __restrict__ X, uint32_t ld, uint32_t n_stages,
    const uint32_t* __restrict__ rowmap, const uint32_t* __restrict__ perm, uint32_t n,
    uint32_t T, const uint2* __restrict__ queue, const unsigned int* __restrict__ qcount,
    unsigned int* __restrict__ qhead, uint32_t* __restrict__ tri,
    unsigned long long* __restrict__ sum, unsigned long long* __restrict__ tile_ticks) {
    __shared__ __attribute__((aligned(16))) BBMfmaSmem<FP4> sm;
    __shared__ uint32_t s_t;
    uint64_t local = 0;
    if (!qcount) {
        uint32_t bi, bj;
        tri_tile(blockIdx.x, T, bi, bj);
        local = bb_mfma_tile<FP4>(sm, X, ld, n_stages, rowmap, perm, n, bi, bj, tri);
    } else {
        for (;;) {
            __syncthreads();
            if (threadIdx.x == 0) s_t = atomicAdd(qhead, 1u);
            __syncthreads();
            const uint32_t t = s_t;
            if (t >= *qcount) break;
            const uint2 tp = queue[t];
            const uint64_t t0 = wall_clock64();
            local += bb_mfma_tile<FP4>(sm, X, ld, n_stages, rowmap, perm, n, tp.x, tp.y, tri);
            if (tile_ticks) {
                __syncthreads();
                if (threadIdx.x == 0) tile_ticks[t] = wall_clock64() - t0;
            }
        }
    }
    local = wave_sum_u64(local);
    if ((threadIdx.x & 63) == 0 && local) atomicAdd(sum, (unsigned long long)local);
}

int launch_bb_mfma_fmt(bool fp4, hipStream_t s, const uint64_t* X, uint32_t ld, uint32_t n_words,
                       const uint32_t* rowmap, const uint32_t* perm, uint32_t n,
                       const uint2* queue, const unsigned int* qcount, unsigned int* qhead,
                       uint32_t grid, uint32_t* tri, unsigned long long* sum,
                       unsigned long long* tile_ticks) {
    const uint32_t n_stages = (n_words + kWordsStage - 1) / kWordsStage;
    if ((uint64_t)(n_stages ? n_stages : 1) * kWordsStage > ld) {
        set_error("bb_mfma: %u stages exceed row stride %u", n_stages, ld);
        return STORM_HIP_ERR_INVALID;
    }
    // Exactness bounds: i32 accumulators below 2^31 bits; f32 below 2^24.
    if (n_words >= (fp4 ? (1u << 18) : (1u << 25))) {
        set_error("bb_mfma: universe of %u words too large for exact %s accumulation", n_words,
                  fp4 ? "f32" : "i32");
        return STORM_HIP_ERR_RANGE;
    }
    const uint32_t T = (n + kTile - 1) / kTile;
    const uint64_t g = queue ? grid : (uint64_t)T * (T + 1) / 2;
    if (g == 0) return STORM_HIP_OK;
    if (fp4)
        hipLaunchKernelGGL(bb_mfma_kernel<true>, dim3((uint32_t)g), dim3(256), 0, s, X, ld,
                           n_stages ? n_stages : 1, rowmap, perm, n, T, queue, qcount, qhead, tri,
                           sum, tile_ticks);
    else
        hipLaunchKernelGGL(bb_mfma_kernel<false>, dim3((uint32_t)g), dim3(256), 0, s, X, ld,
                           n_stages ? n_stages : 1, rowmap, perm, n, T, queue, qcount, qhead, tri,
                           sum, tile_ticks);
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

int launch_bb_dense(uint32_t bb_kernel, hipStream_t s, const uint64_t* X, uint32_t ld,
                    uint32_t n_words, const uint32_t* rowmap, uint32_t n, uint32_t* tri,
                    unsigned long long* sum) {
    if (bb_kernel == STORM_HIP_BB_VALU)
        return launch_bb_valu(s, 128, X, ld, n_words, rowmap, nullptr, n, nullptr, nullptr,
                              nullptr, 0, tri, sum);
    const bool fp4 = bb_kernel == STORM_HIP_BB_MFMA_FP4 ||
                     (bb_kernel == STORM_HIP_BB_AUTO && n_words < (1u << 18));
    return launch_bb_mfma_fmt(fp4, s, X, ld, n_words, rowmap, nullptr, n, nullptr, nullptr,
                              nullptr, 0, tri, sum);
}

int launch_bb_mfma(hipStream_t s, const uint64_t* X, uint32_t ld, uint32_t n_words,
                   const uint32_t* rowmap, const uint32_t* perm, uint32_t n, const uint2* queue,
                   const unsigned int* qcount, unsigned int* qhead, uint32_t grid, uint32_t* tri,
                   unsigned long long* sum, unsigned long long* tile_ticks) {
    return launch_bb_mfma_fmt(false, s, X, ld, n_words, rowmap, perm, n, queue, qcount, qhead,
                              grid, tri, sum, tile_ticks);
}

}  // namespace stormhip
