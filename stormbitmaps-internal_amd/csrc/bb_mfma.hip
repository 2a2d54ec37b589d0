// B x B cell, matrix-core form: exact |Xi n Xj| as an int8 Gram matrix.
//
// Each bit of a row becomes one int8 (0 or 1); |Xi n Xj| = sum_k Xi[k]*Xj[k]
// with 32-bit integer accumulation, so the count is exact for any universe
// below 2^31 bits. The expansion bits -> bytes is done ONCE per workgroup
// stage into LDS and shared by the four waves, which then run
// v_mfma_i32_32x32x32_i8 over it.
//
// Same contract as the VALU form (bb_valu.hip): a 256 x 256 block of pairs per
// workgroup, upper triangle of blocks, optional rowmap/perm indirection.
//
// Operand mapping: lane l of an A (or B) fragment holds 16 int8 values for
// row (l % 32) of its 32-row block, taken from bits [16*(l/32), 16*(l/32)+16)
// of the current 32-bit k-chunk of that row. A and B use the identical
// mapping, so each MFMA sums the products of the same 32 bit positions.
#include "device_common.h"
#include "internal.h"

namespace stormhip {

typedef int8_t i8x16 __attribute__((ext_vector_type(16)));
typedef int32_t i32x16 __attribute__((ext_vector_type(16)));

namespace {
constexpr int kTile = 256;     // pairs block: kTile x kTile
constexpr int kWordsStage = 2; // packed 64-bit words per LDS stage
constexpr int kChunks = kWordsStage * 2;  // 32-bit k-chunks per stage
constexpr int kRowBytes = 32;  // expanded bytes per row per k-chunk

// 4 bits -> 4 bytes of 0/1: nibble * 0x00204081 places bit b at bit 8b+... ;
// the mask keeps bit 0 of every byte.
__device__ inline uint32_t nib_to_bytes(uint32_t x) { return (x * 0x00204081u) & 0x01010101u; }

// One 32-bit chunk -> 32 bytes (byte k = bit k) as 8 dwords.
__device__ inline void expand32(uint32_t d, uint4& lo, uint4& hi) {
    lo.x = nib_to_bytes(d & 0xF);
    lo.y = nib_to_bytes((d >> 4) & 0xF);
    lo.z = nib_to_bytes((d >> 8) & 0xF);
    lo.w = nib_to_bytes((d >> 12) & 0xF);
    hi.x = nib_to_bytes((d >> 16) & 0xF);
    hi.y = nib_to_bytes((d >> 20) & 0xF);
    hi.z = nib_to_bytes((d >> 24) & 0xF);
    hi.w = nib_to_bytes(d >> 28);
}
}  // namespace

__global__ __launch_bounds__(256, 1) void bb_mfma_kernel(
    const uint64_t* __restrict__ X, uint32_t ld, uint32_t n_stages,
    const uint32_t* __restrict__ rowmap, const uint32_t* __restrict__ perm, uint32_t n,
    uint32_t T, const uint32_t* __restrict__ tiles, uint32_t* __restrict__ tri,
    unsigned long long* __restrict__ sum) {
    // Expanded operands of one stage: [chunk][row][32 bytes], A then B.
    __shared__ __attribute__((aligned(16))) uint8_t lds[2][kChunks][kTile][kRowBytes];

    uint32_t bi, bj;
    if (tiles) { bi = tiles[2 * blockIdx.x]; bj = tiles[2 * blockIdx.x + 1]; }
    else       tri_tile(blockIdx.x, T, bi, bj);
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

    i32x16 acc[4][4];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[m][q][v] = 0;

    uint4 ga = pa[0], gb = pb[0];
    for (uint32_t s = 0; s < n_stages; ++s) {
        __syncthreads();   // previous stage's fragments fully read
        {
            const uint32_t da[4] = {ga.x, ga.y, ga.z, ga.w};
            const uint32_t db[4] = {gb.x, gb.y, gb.z, gb.w};
#pragma unroll
            for (int c = 0; c < kChunks; ++c) {
                uint4 lo, hi;
                expand32(da[c], lo, hi);
                *reinterpret_cast<uint4*>(&lds[0][c][t][0]) = lo;
                *reinterpret_cast<uint4*>(&lds[0][c][t][16]) = hi;
                expand32(db[c], lo, hi);
                *reinterpret_cast<uint4*>(&lds[1][c][t][0]) = lo;
                *reinterpret_cast<uint4*>(&lds[1][c][t][16]) = hi;
            }
        }
        __syncthreads();
        if (s + 1 < n_stages) { ga = pa[s + 1]; gb = pb[s + 1]; }
#pragma unroll
        for (int c = 0; c < kChunks; ++c) {
            i8x16 fa[4], fb[4];
#pragma unroll
            for (int m = 0; m < 4; ++m)
                fa[m] = *reinterpret_cast<const i8x16*>(&lds[0][c][wr * 128 + m * 32 + lr][lg * 16]);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                fb[q] = *reinterpret_cast<const i8x16*>(&lds[1][c][wc * 128 + q * 32 + lr][lg * 16]);
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    acc[m][q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[m], fb[q], acc[m][q], 0, 0, 0);
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
    local = wave_sum_u64(local);
    if (lane == 0 && local) atomicAdd(sum, (unsigned long long)local);
}

int launch_bb_mfma(hipStream_t s, const uint64_t* X, uint32_t ld, uint32_t n_words,
                   const uint32_t* rowmap, const uint32_t* perm, uint32_t n,
                   const uint32_t* tiles, uint64_t n_tiles, uint32_t* tri,
                   unsigned long long* sum) {
    const uint32_t n_stages = (n_words + kWordsStage - 1) / kWordsStage;
    if ((uint64_t)(n_stages ? n_stages : 1) * kWordsStage > ld) {
        set_error("bb_mfma: %u stages exceed row stride %u", n_stages, ld);
        return STORM_HIP_ERR_INVALID;
    }
    if (n_words >= (1u << 25)) {  // 2^31 bits: i32 accumulators stay exact below this
        set_error("bb_mfma: universe too large for i32 accumulation");
        return STORM_HIP_ERR_RANGE;
    }
    const uint32_t T = (n + kTile - 1) / kTile;
    const uint64_t grid = tiles ? n_tiles : (uint64_t)T * (T + 1) / 2;
    if (grid == 0) return STORM_HIP_OK;
    hipLaunchKernelGGL(bb_mfma_kernel, dim3((uint32_t)grid), dim3(256), 0, s, X, ld,
                       n_stages ? n_stages : 1, rowmap, perm, n, T, tiles, tri, sum);
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

}  // namespace stormhip
