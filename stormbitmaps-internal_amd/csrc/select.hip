// Device-side selection: per-tile metadata, the analytic cost model of
// kernels/storm_cost.cpp ported to the device, eligibility against the
// representations actually resident, and the per-cell work queues.
#include "device_common.h"
#include "general.h"

namespace stormhip {

// tile_meta (kernels/storm_allpairs.cpp:281-312) plus eligibility/spread
// statistics, one workgroup per row tile of density-sorted rows.
__global__ __launch_bounds__(256) void tile_meta_kernel(DevCorpus D, uint32_t tile,
                                                        TileMeta* __restrict__ out) {
    const uint32_t t = blockIdx.x;
    const uint32_t lo = t * tile, hi = min(D.n, lo + tile);
    uint64_t card = 0, runs = 0, nnz = 0;
    uint32_t first = 0xFFFFFFFFu, last = 0, mxc = 0, mnc = 0xFFFFFFFFu, mxr = 0, mxe = 0, nw = 0;
    uint32_t rank_and = 1, occ_and = 1, bm_and = 1, comp_and = 1, empty_or = 0;
    for (uint32_t k = lo + threadIdx.x; k < hi; k += blockDim.x) {
        const uint32_t r = D.perm[k];
        const uint32_t* m = D.meta + (uint64_t)r * 8;
        card += m[0]; runs += m[1]; nnz += m[3];
        nw = max(nw, m[2]);
        if (m[0]) { first = min(first, m[4]); last = max(last, m[5]); }
        mxc = max(mxc, m[0]); mnc = min(mnc, m[0]); mxr = max(mxr, m[1]);
        mxe = max(mxe, (uint32_t)(D.mk_off[r + 1] - D.mk_off[r]));
        rank_and &= m[6]; occ_and &= m[7];
        bm_and &= D.bm_row[r] != kNoBitmap;
        comp_and &= 2ull * m[0] > D.universe;
        empty_or |= m[0] == 0;
    }
    __shared__ uint64_t s64[3][256];
    __shared__ uint32_t s32[12][256];
    const int x = threadIdx.x;
    s64[0][x] = card; s64[1][x] = runs; s64[2][x] = nnz;
    s32[0][x] = first; s32[1][x] = last; s32[2][x] = mxc; s32[3][x] = mnc; s32[4][x] = mxr;
    s32[5][x] = mxe; s32[6][x] = nw; s32[7][x] = rank_and; s32[8][x] = occ_and;
    s32[9][x] = bm_and; s32[10][x] = comp_and; s32[11][x] = empty_or;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (x < st) {
            for (int q = 0; q < 3; ++q) s64[q][x] += s64[q][x + st];
            s32[0][x] = min(s32[0][x], s32[0][x + st]);
            s32[1][x] = max(s32[1][x], s32[1][x + st]);
            s32[2][x] = max(s32[2][x], s32[2][x + st]);
            s32[3][x] = min(s32[3][x], s32[3][x + st]);
            s32[4][x] = max(s32[4][x], s32[4][x + st]);
            s32[5][x] = max(s32[5][x], s32[5][x + st]);
            s32[6][x] = max(s32[6][x], s32[6][x + st]);
            for (int q = 7; q <= 10; ++q) s32[q][x] &= s32[q][x + st];
            s32[11][x] |= s32[11][x + st];
        }
        __syncthreads();
    }
    if (x == 0) {
        const uint32_t cnt = hi > lo ? hi - lo : 1;
        TileMeta m;
        m.card = (uint32_t)(s64[0][0] / cnt);
        m.runs = (uint32_t)(s64[1][0] / cnt);
        m.nnz = (uint32_t)(s64[2][0] / cnt);
        m.nw = s32[6][0];
        m.first = s32[0][0] == 0xFFFFFFFFu ? 0 : s32[0][0];
        m.last = s32[1][0];
        m.has_rank = hi > lo && s32[7][0];
        m.has_occ = hi > lo && s32[8][0];
        m.all_bitmap = s32[9][0];
        m.all_comp = s32[10][0];
        m.any_empty = s32[11][0];
        m.max_card = s32[2][0];
        m.min_card = s32[3][0];
        m.max_runs = s32[4][0];
        m.max_ewah = s32[5][0];
        m.rows = hi - lo;
        out[t] = m;
    }
}

// ---- the analytic model (kernels/storm_cost.cpp:195-465) -----------------
__host__ __device__ inline uint32_t bit_length(uint32_t x) {   // 32 - clz(x)
    uint32_t b = 0;
    while (x) { ++b; x >>= 1; }
    return b;
}

__host__ __device__ inline double bb_expected_bins(const TileMeta& a, const TileMeta& b) {
    const double bins = fmax(1.0, (double)max(a.nw, b.nw) / 8.0);
    const double fa = fmin(1.0, (double)a.nnz / (double)max(1u, a.nw));
    const double fb = fmin(1.0, (double)b.nnz / (double)max(1u, b.nw));
    const double pa = fmin(1.0, 8.0 * fa), pb = fmin(1.0, 8.0 * fb);
    return fmax(1.0, bins * pa * pb * 8.0);
}

__host__ __device__ double work_units(int p, const TileMeta& a, const TileMeta& b, double margin) {
    const double sa = a.card, sb = b.card, ra = a.runs, rb = b.runs;
    const double wa = 2.0 * ra + 1.0 + (double)a.nnz, wb = 2.0 * rb + 1.0 + (double)b.nnz;
    const bool a_sparse = sa <= sb;
    const double s_card = a_sparse ? sa : sb;
    const double d_runs = a_sparse ? rb : ra;
    const double d_ewah = a_sparse ? wb : wa;
    switch (p) {
        case STORM_HIP_CELL_BB: return bb_expected_bins(a, b);
        case STORM_HIP_CELL_BS: return s_card;
        case STORM_HIP_CELL_BR: {
            const double dr = a_sparse ? rb : ra, sr = a_sparse ? ra : rb;
            return (margin > 0.0 && dr * margin <= sr) ? dr : sr;
        }
        case STORM_HIP_CELL_BW: {
            const double dr = a_sparse ? rb : ra, sr = a_sparse ? ra : rb;
            const bool sw = margin > 0.0 && dr * margin <= sr;
            return sw ? (2.0 * dr + 1.0 + (a_sparse ? b.nnz : a.nnz))
                      : (2.0 * sr + 1.0 + (a_sparse ? a.nnz : b.nnz));
        }
        case STORM_HIP_CELL_SS: {
            const uint32_t mx = (uint32_t)fmax(2.0, fmax(sa, sb));
            return fmin(sa, sb) * (double)bit_length(mx);
        }
        case STORM_HIP_CELL_SR: return s_card + d_runs;
        case STORM_HIP_CELL_SW: return s_card + d_ewah;
        case STORM_HIP_CELL_RR: return ra + rb;
        case STORM_HIP_CELL_RW: return fmin(ra + wb, rb + wa);
        case STORM_HIP_CELL_WW: return wa + wb;
        default: return 0.0;
    }
}

__host__ __device__ double predict(const DevModel& m, int p, const TileMeta& a, const TileMeta& b) {
    if (p == STORM_HIP_CELL_EMPTY) return 0.0;
    if (p == STORM_HIP_CELL_BB && (m.bb_zonemap == 0.0 || !(a.has_occ && b.has_occ)))
        return m.ns_fixed + m.ns_per_unit[p] * (double)max(a.nw, b.nw);
    double c = m.ns_fixed + m.ns_per_unit[p] * work_units(p, a, b, m.br_margin) *
                                (p == STORM_HIP_CELL_RR ? m.rr_scale : 1.0);
    if (p == STORM_HIP_CELL_BS && m.ns_probe_cold > m.ns_probe_hot) {
        const double touches = a.card <= b.card ? (double)a.nnz : (double)b.nnz;
        c += (m.ns_probe_cold - m.ns_probe_hot) * m.probe_scale * touches;
    }
    if (p == STORM_HIP_CELL_BB) c += m.ns_per_occ_word * fmax(1.0, (double)max(a.nw, b.nw) / 512.0);
    return c;
}

// select_pairing over the reference's candidate set (BB seeded, uncalibrated
// rates excluded), storm_cost.cpp:422-465. Returns EMPTY on its emptiness proof.
__device__ int select_pairing(const DevModel& m, const TileMeta& a, const TileMeta& b, double* cost) {
    if (a.card == 0 || b.card == 0 || a.last < b.first || b.last < a.first) {
        *cost = 0.0;
        return STORM_HIP_CELL_EMPTY;
    }
    int best = STORM_HIP_CELL_BB;
    double bc = predict(m, best, a, b);
    for (int p = 0; p < 10; ++p) {
        if (m.ns_per_unit[p] <= 0.0) continue;
        const double c = predict(m, p, a, b);
        if (c < bc) { bc = c; best = p; }
    }
    *cost = bc;
    return best;
}

__device__ inline int cheapest_eligible(const DevModel& m, const TileMeta& a, const TileMeta& b,
                                        uint32_t elig) {
    int best = -1;
    double bc = 0.0;
    for (int p = 0; p < 10; ++p) {
        if (!(elig >> p & 1)) continue;
        const double c = predict(m, p, a, b);
        if (best < 0 || c < bc) { bc = c; best = p; }
    }
    return best < 0 ? STORM_HIP_CELL_SS : best;
}

// Cells the learned selector may rank on a tile: eligible cells whose
// analytic prediction is within kLearnedPrune times the cheapest eligible
// one. The same set is timed when training samples are drawn, so the network
// never ranks a cell it had no chance to see on such a tile.
__host__ __device__ uint32_t learned_candidates(const DevModel& m, const TileMeta& a,
                                                const TileMeta& b, uint32_t elig) {
    double c[10];
    double lo = -1.0;
    for (int p = 0; p < 10; ++p) {
        if (!(elig >> p & 1)) continue;
        c[p] = predict(m, p, a, b);
        if (lo < 0.0 || c[p] < lo) lo = c[p];
    }
    uint32_t mask = 0;
    for (int p = 0; p < 10; ++p)
        if ((elig >> p & 1) && c[p] <= kLearnedPrune * lo + 1.0) mask |= 1u << p;
    return mask;
}

// One thread per tile pair: decide, then append to that cell's queue.
// policy: 1 analytic, 2 learned (scores given), 3 fixed.
__global__ __launch_bounds__(256) void select_kernel(
    const TileMeta* __restrict__ tm, uint32_t T, uint64_t P, DevModel M, uint32_t reps,
    uint32_t policy, uint32_t fixed_cell, const float* __restrict__ learned_cost,
    const uint8_t* __restrict__ learned_ok, uint8_t* __restrict__ decisions,
    QItem* __restrict__ queue, uint64_t qstride, uint32_t inst, bool split, double thr,
    uint32_t bb_ksplit, bool bb_wide, uint64_t explore_seed, uint32_t explore_share,
    const uint32_t* __restrict__ sel, DevCounters* __restrict__ ctr) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P) return;
    uint32_t bi, bj;
    tri_tile((uint32_t)t, T, bi, bj);
    // Band and cross calls queue only the tile pairs of their range.
    if (bi < sel[0] || bi >= sel[1] || bj < sel[2] || bj >= sel[3]) {
        decisions[t] = 0xFF;
        return;
    }
    const TileMeta a = tm[bi], b = tm[bj];
    const uint32_t elig = eligible_mask(a, b, reps);
    int p;
    bool fallback = false;
    // Thresholded calls: a tile whose length bound is below the threshold
    // holds no reportable pair.
    if (tile_empty(a, b) || (thr > 0.0 && length_bound(a, b) < thr)) {
        p = STORM_HIP_CELL_EMPTY;
    } else if (policy == kPolicyExplore &&
               (splitmix_mix(explore_seed ^ (t * 0xD1B54A32D192ED03ull)) >> 40) <
                   (uint64_t)explore_share) {
        // This tile explores: a cell drawn among the learned selector's
        // candidates; the other tiles of an exploration call run the analytic
        // choice (below) and are labelled with its cost.
        const uint32_t cand = learned_candidates(M, a, b, elig);
        const uint32_t k = __popc(cand);
        if (k == 0) {
            p = cheapest_eligible(M, a, b, elig);
        } else {
            uint32_t r = (uint32_t)(splitmix_mix(explore_seed + (t + 1) * kGolden) % k);
            p = 0;
            for (uint32_t m = cand;; m &= m - 1) {
                if (r-- == 0) { p = __ffs(m) - 1; break; }
            }
        }
    } else if (policy == STORM_HIP_POLICY_FIXED) {
        p = (int)fixed_cell;
        if (p == STORM_HIP_CELL_EMPTY || !(elig >> p & 1)) { p = cheapest_eligible(M, a, b, elig); fallback = true; }
    } else if (policy == STORM_HIP_POLICY_LEARNED && learned_ok && learned_ok[t]) {
        const uint32_t cand = learned_candidates(M, a, b, elig);
        int best = -1;
        float bc = 0.f;
        for (int q = 0; q < 10; ++q) {
            if (!(cand >> q & 1)) continue;
            const float c = learned_cost[t * 10 + q];
            if (best < 0 || c < bc) { bc = c; best = q; }
        }
        p = best < 0 ? cheapest_eligible(M, a, b, elig) : best;
        // Trust region: the analytic choice stands unless the network predicts
        // the other cell faster by more than kLearnedMargin (log2 units).
        double ca;
        int pa = select_pairing(M, a, b, &ca);
        if (pa == STORM_HIP_CELL_EMPTY) pa = STORM_HIP_CELL_BB;
        if (!(elig >> pa & 1)) pa = cheapest_eligible(M, a, b, elig);
        if (best >= 0 && pa != best && (cand >> pa & 1) &&
            learned_cost[t * 10 + pa] - bc <= kLearnedMargin)
            p = pa;
    } else {
        if (policy == STORM_HIP_POLICY_LEARNED) fallback = true;
        double c;
        p = select_pairing(M, a, b, &c);
        // The reference never skips a whole tile on the mean-based proof
        // (storm_allpairs.cpp:592): it runs B x B and lets the span test act.
        if (p == STORM_HIP_CELL_EMPTY) p = STORM_HIP_CELL_BB;
        if (!(elig >> p & 1)) p = cheapest_eligible(M, a, b, elig);
    }
    decisions[t] = (uint8_t)p;
    if (split && p == STORM_HIP_CELL_BS) {
        // B x S: stripes of kSplitRows probed rows against the whole other
        // side (read through its slices).
        const bool pa = bs_probe_a(a, b);
        const uint32_t ns = ((pa ? a.rows : b.rows) + kSplitRows - 1) / kSplitRows;
        uint32_t slot = atomicAdd(&ctr->qcount[p], ns);
        for (uint32_t x = 0; x < ns; ++x)
            queue[(uint64_t)p * qstride + slot++] =
                QItem{inst, bi, bj, pa ? make_sub(kSplitRows, x, kWholeSide)
                                       : make_sub(kSplitRows, kWholeSide, x)};
    } else if (split && p != STORM_HIP_CELL_BB && p != STORM_HIP_CELL_EMPTY) {
        // Per-pair cells: the tile as kSplitRows x kSplitRows blocks (upper
        // triangle of blocks on a diagonal tile), each its own work item.
        const uint32_t ni = (a.rows + kSplitRows - 1) / kSplitRows;
        const uint32_t nj = (b.rows + kSplitRows - 1) / kSplitRows;
        const uint32_t count = bi == bj ? ni * (ni + 1) / 2 : ni * nj;
        uint32_t slot = atomicAdd(&ctr->qcount[p], count);
        for (uint32_t x = 0; x < ni; ++x)
            for (uint32_t y = bi == bj ? x : 0; y < nj; ++y)
                queue[(uint64_t)p * qstride + slot++] =
                    QItem{inst, bi, bj, make_sub(kSplitRows, x, y)};
    } else if (p == STORM_HIP_CELL_BB) {
        const uint32_t q = bb_wide ? kQueueBBWide : STORM_HIP_CELL_BB;
        const uint32_t ks = bb_ksplit > 1 ? bb_ksplit : 1;
        uint32_t slot = atomicAdd(&ctr->qcount[q], ks);
        for (uint32_t x = 0; x < ks; ++x)
            queue[(uint64_t)q * qstride + slot++] =
                QItem{inst, bi, bj, ks > 1 ? (x | (ks << 16)) : 0u};
    } else {
        const uint32_t slot = atomicAdd(&ctr->qcount[p], 1u);
        queue[(uint64_t)p * qstride + slot] = QItem{inst, bi, bj, 0};
    }
    atomicAdd(&ctr->cell_tiles[p], 1ull);
    if (p == STORM_HIP_CELL_BB) {
        // B x B tiles skip the per-pair span test (same cost either way);
        // their pairs are counted here.
        const uint64_t ni = a.rows, nj = b.rows;
        atomicAdd(&ctr->cell_pairs[p], (unsigned long long)(bi == bj ? ni * (ni - 1) / 2 : ni * nj));
    }
    if (fallback) atomicAdd(&ctr->fallback_tiles, 1ull);
}

// Host view of the work function, for calibration.
double host_work_units(int p, const TileMeta& a, const TileMeta& b, double margin) {
    if (p == STORM_HIP_CELL_BB) return (double)std::max(a.nw, b.nw);   // device B x B is Theta(m)
    return work_units(p, a, b, margin);
}

// rowmap[pos] = bitmap row of the row at sorted position pos.
__global__ void compose_rowmap_kernel(const uint32_t* __restrict__ order,
                                      const uint32_t* __restrict__ bm_row, uint32_t npad,
                                      uint32_t* __restrict__ rowmap) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < npad) rowmap[k] = bm_row[order[k]];
}

// Host-side selection entry: tile metas + select into queues.
int launch_tile_meta(hipStream_t s, const DevCorpus& D, uint32_t tile, uint32_t T, TileMeta* out) {
    if (T == 0) return STORM_HIP_OK;
    hipLaunchKernelGGL(tile_meta_kernel, dim3(T), dim3(256), 0, s, D, tile, out);
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

int launch_select(hipStream_t s, const TileMeta* tm, uint32_t T, const DevModel& M, uint32_t reps,
                  uint32_t policy, uint32_t fixed_cell, const float* learned_cost,
                  const uint8_t* learned_ok, uint8_t* decisions, QItem* queue, uint64_t qstride,
                  uint32_t inst, uint32_t tile, double thr, uint32_t bb_ksplit,
                  bool bb_wide, uint64_t explore_seed, uint32_t explore_share,
                  const uint32_t* sel, DevCounters* ctr) {
    const uint64_t P = (uint64_t)T * (T + 1) / 2;
    if (P == 0) return STORM_HIP_OK;
    hipLaunchKernelGGL(select_kernel, dim3((uint32_t)((P + 255) / 256)), dim3(256), 0, s, tm, T, P,
                       M, reps, policy, fixed_cell, learned_cost, learned_ok, decisions, queue,
                       qstride, inst, tile > kSplitRows, thr, bb_ksplit, bb_wide, explore_seed,
                       explore_share, sel, ctr);
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

int launch_compose_rowmap(hipStream_t s, const uint32_t* order, const uint32_t* bm_row,
                          uint32_t npad, uint32_t* rowmap) {
    hipLaunchKernelGGL(compose_rowmap_kernel, dim3((npad + 255) / 256), dim3(256), 0, s, order,
                       bm_row, npad, rowmap);
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

}  // namespace stormhip
