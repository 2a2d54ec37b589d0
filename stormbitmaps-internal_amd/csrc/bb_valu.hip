// B x B cell, VALU form: exact |Xi n Xj| for a tile of row pairs as
// AND + v_bcnt_u32_b32 popcount accumulate over LDS-staged row panels.
//
// Replaces the reference's bb_dense / simd_and_popcnt (kernels/cell_bb.cpp:402,
// kernels/storm_simd.h:144-203) and libalgebra's Harley-Seal kernel behind
// STORM_contig_pairw_intersect_cardinality (storm.cpp:1232-1240) -- but as a
// Gram-matrix tile: each workgroup owns a BM x BM block of pairs, so a row
// panel is read once per block instead of once per pair.
//
// Layout: rows are `ld` u64 words apart (ld a multiple of 64, zero padded);
// `rowmap` (optional) maps a tile-row index to a physical row of X; `perm`
// (optional) maps a tile-row index to the ORIGINAL row id used for output.
// Tiles come from the grid (upper triangle of the T x T block grid) or, in
// queue mode, from a device queue drained by persistent workgroups.
#include "device_common.h"
#include "general.h"

namespace stormhip {

template <int BM>
struct BBValuSmem {
    static constexpr int KC = 16;   // words per LDS stage
    uint64_t A[2][KC][BM];
    uint64_t B[2][KC][BM];
};

// THR (thresholded queue launches): pairs go through the output of the corpus
// view `Dp` instead of `tri`.
template <int BM, bool THR = false>
__device__ uint64_t bb_valu_tile(BBValuSmem<BM>& sm, const uint64_t* __restrict__ X, uint32_t ld,
                                 uint32_t kwords, const uint32_t* __restrict__ rowmap,
                                 const uint32_t* __restrict__ perm, uint32_t n, uint32_t bi,
                                 uint32_t bj, uint32_t* __restrict__ tri,
                                 const DevCorpus* __restrict__ Dp = nullptr) {
    constexpr int KC = BBValuSmem<BM>::KC;
    constexpr int TM = BM / 16;          // rows (and cols) per thread
    constexpr int WPT = BM * KC / 256;   // words each thread loads per panel
    constexpr int TPR = KC / WPT;        // threads per panel row
    const uint32_t i0 = bi * BM, j0 = bj * BM;
    const int t = threadIdx.x, ty = t >> 4, tx = t & 15;
    const int lrow = t / TPR, lseg = t % TPR;
    const uint64_t ra = rowmap ? rowmap[i0 + lrow] : (uint64_t)(i0 + lrow);
    const uint64_t rb = rowmap ? rowmap[j0 + lrow] : (uint64_t)(j0 + lrow);
    const uint4* pa = reinterpret_cast<const uint4*>(X + ra * ld + lseg * WPT);
    const uint4* pb = reinterpret_cast<const uint4*>(X + rb * ld + lseg * WPT);

    uint4 ga[WPT / 2], gb[WPT / 2];
    auto gload = [&](uint32_t k0) {
#pragma unroll
        for (int q = 0; q < WPT / 2; ++q) {
            ga[q] = pa[k0 / 2 + q];
            gb[q] = pb[k0 / 2 + q];
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < WPT / 2; ++q) {
            sm.A[buf][lseg * WPT + 2 * q][lrow]     = ((uint64_t)ga[q].y << 32) | ga[q].x;
            sm.A[buf][lseg * WPT + 2 * q + 1][lrow] = ((uint64_t)ga[q].w << 32) | ga[q].z;
            sm.B[buf][lseg * WPT + 2 * q][lrow]     = ((uint64_t)gb[q].y << 32) | gb[q].x;
            sm.B[buf][lseg * WPT + 2 * q + 1][lrow] = ((uint64_t)gb[q].w << 32) | gb[q].z;
        }
    };

    uint32_t acc[TM][TM];
#pragma unroll
    for (int r = 0; r < TM; ++r)
#pragma unroll
        for (int c = 0; c < TM; ++c) acc[r][c] = 0;

    const uint32_t nk = kwords / KC;
    __syncthreads();   // LDS free (previous tile of a persistent workgroup)
    gload(0);
    sstore(0);
    __syncthreads();
    for (uint32_t kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) gload((kt + 1) * KC);
#pragma unroll 4
        for (int k = 0; k < KC; ++k) {
            uint64_t a[TM], b[TM];
#pragma unroll
            for (int r = 0; r < TM; ++r) a[r] = sm.A[buf][k][ty + 16 * r];
#pragma unroll
            for (int c = 0; c < TM; ++c) b[c] = sm.B[buf][k][tx + 16 * c];
#pragma unroll
            for (int r = 0; r < TM; ++r)
#pragma unroll
                for (int c = 0; c < TM; ++c) {
                    const uint64_t x = a[r] & b[c];
                    acc[r][c] += __popc((uint32_t)x) + __popc((uint32_t)(x >> 32));
                }
        }
        if (kt + 1 < nk) sstore(buf ^ 1);
        __syncthreads();
    }

    uint64_t local = 0;
#pragma unroll
    for (int r = 0; r < TM; ++r) {
        const uint32_t i = i0 + ty + 16 * r;
        if (i >= n) continue;
        const uint32_t oi = perm ? perm[i] : i;
#pragma unroll
        for (int c = 0; c < TM; ++c) {
            const uint32_t j = j0 + tx + 16 * c;
            if (j >= n || j <= i) continue;
            const uint32_t v = acc[r][c];
            local += v;
            if constexpr (THR) {
                emit_pair(*Dp, oi, perm ? perm[j] : j, v);
            } else if (tri) {
                const uint32_t oj = perm ? perm[j] : j;
                tri[oi < oj ? tri_index(oi, oj, n) : tri_index(oj, oi, n)] = v;
            }
        }
    }
    return local;
}

template <int BM, bool THR = false>
__global__ __launch_bounds__(256, 2) void bb_valu_kernel(
    const uint64_t* __restrict__ X, uint32_t ld, uint32_t kwords,
    const uint32_t* __restrict__ rowmap, const uint32_t* __restrict__ perm, uint32_t n,
    uint32_t T, const DevCorpus* __restrict__ Ds, const QItem* __restrict__ queue,
    const unsigned int* __restrict__ qcount, unsigned int* __restrict__ qhead,
    uint32_t* __restrict__ tri, unsigned long long* __restrict__ sum,
    unsigned long long* __restrict__ tile_ticks) {
    __shared__ BBValuSmem<BM> sm;
    __shared__ uint32_t s_t;
    if (THR && !qcount) return;   // thresholded instances run in queue mode only
    if (!qcount) {
        uint32_t bi, bj;
        tri_tile(blockIdx.x, T, bi, bj);
        uint64_t local = bb_valu_tile<BM>(sm, X, ld, kwords, rowmap, perm, n, bi, bj, tri);
        local = wave_sum_u64(local);
        if ((threadIdx.x & 63) == 0 && local) atomicAdd(sum, (unsigned long long)local);
        return;
    }
    // Queue mode: tiles of any corpus of the call, each with its own arrays.
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_t = atomicAdd(qhead, 1u);
        __syncthreads();
        const uint32_t t = s_t;
        if (t >= *qcount) break;
        const QItem e = queue[t];
        const DevCorpus D = Ds[e.inst];
        const uint64_t t0 = wall_clock64();
        // A position-packed triangle takes the tile's positions as they are.
        const uint32_t* out_perm = (!THR && D.tri_pos) ? nullptr : D.perm;
        uint64_t local = bb_valu_tile<BM, THR>(sm, D.bitmap, D.ld, (D.nw + 15) / 16 * 16,
                                               D.rowmap, out_perm, D.n, e.bi, e.bj, D.tri,
                                               &Ds[e.inst]);
        local = wave_sum_u64(local);
        if ((threadIdx.x & 63) == 0 && local) atomicAdd(D.sum, (unsigned long long)local);
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint64_t now = wall_clock64();
            if (tile_ticks) tile_ticks[t] = now - t0;
            if (D.finish) atomicMax(D.finish, (unsigned long long)now);
        }
    }
}

// Host launcher. queue == nullptr: all T(T+1)/2 upper-triangle tiles of BM
// rows, one workgroup each; otherwise `grid` persistent workgroups drain the
// device queue (qcount/qhead on the device).
int launch_bb_valu(hipStream_t s, int bm, const uint64_t* X, uint32_t ld, uint32_t n_words,
                   const uint32_t* rowmap, const uint32_t* perm, uint32_t n, uint32_t* tri,
                   unsigned long long* sum) {
    const uint32_t kwords = (n_words + 15) / 16 * 16;
    if (kwords > ld) { set_error("bb_valu: kwords %u > ld %u", kwords, ld); return STORM_HIP_ERR_INVALID; }
    const uint32_t T = (n + bm - 1) / bm;
    const uint64_t g = (uint64_t)T * (T + 1) / 2;
    if (g == 0) return STORM_HIP_OK;
    if (bm == 128) {
        hipLaunchKernelGGL(bb_valu_kernel<128>, dim3((uint32_t)g), dim3(256), 0, s, X, ld, kwords,
                           rowmap, perm, n, T, nullptr, nullptr, nullptr, nullptr, tri, sum, nullptr);
    } else if (bm == 64) {
        hipLaunchKernelGGL(bb_valu_kernel<64>, dim3((uint32_t)g), dim3(256), 0, s, X, ld, kwords,
                           rowmap, perm, n, T, nullptr, nullptr, nullptr, nullptr, tri, sum, nullptr);
    } else {
        set_error("bb_valu: unsupported tile %d", bm);
        return STORM_HIP_ERR_INVALID;
    }
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

// Resident workgroups per CU of the VALU kernel for a tile width.
uint32_t bb_valu_blocks_per_cu(int bm) {
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
    if (bm == 128) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, bb_valu_kernel<128>, 256, 0);
    if (bm == 64) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, bb_valu_kernel<64>, 256, 0);
    return (e == hipSuccess && nb > 0) ? (uint32_t)nb : 2;
}

// Queue mode: `grid` persistent workgroups drain a device queue of tiles of
// the corpora in Ds (every corpus's rows padded so ld covers its words).
int launch_bb_valu_queue(hipStream_t s, int bm, const DevCorpus* Ds, const QItem* queue,
                         const unsigned int* qcount, unsigned int* qhead, uint32_t grid,
                         unsigned long long* tile_ticks, bool thresholded) {
    if (grid == 0) return STORM_HIP_OK;
    if (bm == 128) {
        if (thresholded)
            hipLaunchKernelGGL((bb_valu_kernel<128, true>), dim3(grid), dim3(256), 0, s, nullptr,
                               0u, 0u, nullptr, nullptr, 0u, 0u, Ds, queue, qcount, qhead,
                               nullptr, nullptr, tile_ticks);
        else
            hipLaunchKernelGGL(bb_valu_kernel<128>, dim3(grid), dim3(256), 0, s, nullptr, 0u, 0u,
                               nullptr, nullptr, 0u, 0u, Ds, queue, qcount, qhead, nullptr,
                               nullptr, tile_ticks);
    } else if (bm == 64) {
        if (thresholded)
            hipLaunchKernelGGL((bb_valu_kernel<64, true>), dim3(grid), dim3(256), 0, s, nullptr,
                               0u, 0u, nullptr, nullptr, 0u, 0u, Ds, queue, qcount, qhead,
                               nullptr, nullptr, tile_ticks);
        else
            hipLaunchKernelGGL(bb_valu_kernel<64>, dim3(grid), dim3(256), 0, s, nullptr, 0u, 0u,
                               nullptr, nullptr, 0u, 0u, Ds, queue, qcount, qhead, nullptr,
                               nullptr, tile_ticks);
    } else {
        set_error("bb_valu: unsupported tile %d", bm);
        return STORM_HIP_ERR_INVALID;
    }
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

}  // namespace stormhip
