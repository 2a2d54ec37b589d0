// Device-side selection: per-tile metadata, the analytic cost model of
// kernels/storm_cost.cpp ported to the device, eligibility against the
// representations actually resident, and the per-cell work queues.
#include "device_common.h"
#include "general.h"
#include "cost_model.h"

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
        // max / mean cardinality (the refine gate's spread, mean in double).
        const double mean = hi > lo ? (double)s64[0][0] / (double)(hi - lo) : 0.0;
        m.spread = mean > 0.0 ? (double)m.max_card / mean : 1.0;
        out[t] = m;
    }
    // The tile's zone map: the OR of its rows' maps.
    if (D.zone && D.tzone)
        for (uint32_t w = threadIdx.x; w < D.zone_ld; w += blockDim.x) {
            uint64_t v = 0;
            for (uint32_t k = lo; k < hi; ++k) v |= D.zone[(uint64_t)D.perm[k] * D.zone_ld + w];
            D.tzone[(uint64_t)t * D.zone_ld + w] = v;
        }
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

// The tile's two best eligible candidates by predicted cost (tile_top2,
// kernels/storm_allpairs.cpp:330-347, over the cells the tile can run):
// first is select_pairing's choice (EMPTY -> B x B, then the cheapest
// eligible cell when that one is not resident), second the cheapest other
// eligible cell with a calibrated rate (first when there is none); ratio =
// predicted second / first.
__device__ void tile_top2(const DevModel& M, const TileMeta& a, const TileMeta& b, uint32_t elig,
                          int* first, int* second, double* ratio) {
    double c1;
    int p = select_pairing(M, a, b, &c1);
    // The mean-based emptiness proof: B x B, no runner-up (tile_top2 returns
    // Empty for both; the driver runs B x B and lets the span test act).
    const bool proof = p == STORM_HIP_CELL_EMPTY;
    if (proof) p = STORM_HIP_CELL_BB;
    if (!(elig >> p & 1)) p = cheapest_eligible(M, a, b, elig);
    c1 = predict(M, p, a, b);
    int q = p;
    double c2 = 0.0;
    for (int k = 0; k < 10 && !proof; ++k) {
        if (k == p || !(elig >> k & 1) || M.ns_per_unit[k] <= 0.0) continue;
        const double c = predict(M, k, a, b);
        if (q == p || c < c2) { q = k; c2 = c; }
    }
    *first = p;
    *second = q;
    *ratio = (q != p && c1 > 0.0) ? c2 / c1 : 1e300;
}

__device__ inline uint32_t probe_class(int first, int second, double ratio) {
    const uint32_t bucket = ratio < 1.25 ? 0u : ratio < 2.0 ? 1u : ratio < 4.0 ? 2u : 3u;
    return ((uint32_t)first * 10u + (uint32_t)second) * kProbeBuckets + bucket;
}

// Append the work items of tile (bi, bj) run by cell p for instance inst.
__device__ void enqueue_tile(int p, uint32_t inst, uint32_t bi, uint32_t bj, const TileMeta& a,
                             const TileMeta& b, QItem* __restrict__ queue, uint64_t qstride,
                             bool split, uint32_t bb_ksplit, bool bb_wide,
                             DevCounters* __restrict__ ctr) {
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
}

// Append tile (bi, bj) to the refine queue (blocks of kSplitRows when the
// tile is split), its candidates in the items' inst.
__device__ void enqueue_refine(int first, int second, uint32_t inst, uint32_t bi, uint32_t bj,
                               const TileMeta& a, const TileMeta& b, QItem* __restrict__ queue,
                               uint64_t qstride, bool split, DevCounters* __restrict__ ctr) {
    const uint32_t ni = (a.rows + kSplitRows - 1) / kSplitRows;
    const uint32_t nj = (b.rows + kSplitRows - 1) / kSplitRows;
    const uint32_t blocks = split ? (bi == bj ? ni * (ni + 1) / 2 : ni * nj) : 1u;
    const uint32_t tag = inst | ((uint32_t)first << kRefineShift) |
                         ((uint32_t)second << (kRefineShift + 4));
    uint32_t slot = atomicAdd(&ctr->qcount[kQueueRefine], blocks);
    if (!split) {
        queue[(uint64_t)kQueueRefine * qstride + slot] = QItem{tag, bi, bj, 0};
        return;
    }
    for (uint32_t x = 0; x < ni; ++x)
        for (uint32_t y = bi == bj ? x : 0; y < nj; ++y)
            queue[(uint64_t)kQueueRefine * qstride + slot++] =
                QItem{tag, bi, bj, make_sub(kSplitRows, x, y)};
}

// Decide tile pair (bi, bj) = row-major id t and append it to its cell's
// queue. policy: 1 analytic, 2 learned (scores given), 3 fixed, 6 refine, 7
// per pair, and the internal exploration and probe passes. With defer_bs a
// split B x S choice is recorded and counted but left to the caller to queue
// (*deferred set), which may pair it with its neighbour into wide stripes.
__device__ void select_tile(
    uint64_t t, uint32_t bi, uint32_t bj, const TileMeta* __restrict__ tm, const DevModel& M,
    uint32_t reps, uint32_t policy, uint32_t fixed_cell, const float* __restrict__ learned_cost,
    const uint8_t* __restrict__ learned_ok, uint8_t* __restrict__ decisions,
    QItem* __restrict__ queue, uint64_t qstride, uint32_t inst, bool split, double thr,
    uint32_t bb_ksplit, bool bb_wide, uint64_t explore_seed, uint32_t explore_share,
    const uint32_t* __restrict__ sel, DevCounters* __restrict__ ctr, uint16_t* __restrict__ probe,
    ProbeAcc* __restrict__ acc, uint32_t shadow_inst, bool defer_bs, bool* deferred,
    const uint64_t* __restrict__ tzone, uint32_t zone_ld) {
    // Band and cross calls queue only the tile pairs of their range.
    if (bi < sel[0] || bi >= sel[1] || bj < sel[2] || bj >= sel[3]) {
        decisions[t] = 0xFF;
        if (probe) probe[t] = 0;
        return;
    }
    if (policy == kPolicyProbe2) {
        // Commit a pending tile to the candidate its class measured faster
        // (both ran on the class's sampled tiles).
        const uint16_t st = probe[t];
        if (!(st & kProbePending)) return;
        const int first = st & 15, second = (st >> 4) & 15;
        const uint32_t cls = ((uint32_t)first * 10u + (uint32_t)second) * kProbeBuckets +
                             ((st >> 8) & 15);
        const unsigned long long na = acc->ns[0][cls], nb = acc->ns[1][cls];
        const int p = (nb > 0 && na > 0 && nb < na) ? second : first;
        if (p == second) atomicAdd(&acc->switched, 1ull);
        const TileMeta a = tm[bi], b = tm[bj];
        decisions[t] = (uint8_t)p;
        enqueue_tile(p, inst, bi, bj, a, b, queue, qstride, split, bb_ksplit, bb_wide, ctr);
        atomicAdd(&ctr->cell_tiles[p], 1ull);
        if (p == STORM_HIP_CELL_BB)
            atomicAdd(&ctr->cell_pairs[p],
                      (unsigned long long)(bi == bj ? (uint64_t)a.rows * (a.rows - 1) / 2
                                                    : (uint64_t)a.rows * b.rows));
        return;
    }
    const TileMeta a = tm[bi], b = tm[bj];
    const uint32_t elig = eligible_mask(a, b, reps);
    int p;
    bool fallback = false;
    if (probe) probe[t] = 0;
    // Thresholded calls: a tile whose length bound is below the threshold
    // holds no reportable pair.
    if (tile_empty(a, b) || (thr > 0.0 && length_bound(a, b) < thr) ||
        (tzone && tile_zone_shared(tzone, zone_ld, bi, bj) == 0)) {
        p = STORM_HIP_CELL_EMPTY;
    } else if (policy == STORM_HIP_POLICY_PER_PAIR) {
        // Every pair of the tile selects for itself.
        decisions[t] = (uint8_t)kPerPairCell;
        enqueue_refine(kPerPairCell, kPerPairCell, inst, bi, bj, a, b, queue, qstride, split, ctr);
        return;
    } else if (policy == kPolicyProbe1 || policy == STORM_HIP_POLICY_REFINE) {
        int first, second;
        double ratio;
        tile_top2(M, a, b, elig, &first, &second, &ratio);
        p = first;
        if (policy == STORM_HIP_POLICY_REFINE) {
            // Refine when the tile's rows are heterogeneous (max / mean
            // cardinality, either side) and the runner-up is close.
            const double het = fmax(a.spread, b.spread);
            if (second != first && het >= kRefineHet && ratio <= kRefineRatio) {
                decisions[t] = (uint8_t)first;
                enqueue_refine(first, second, inst, bi, bj, a, b, queue, qstride, split, ctr);
                atomicAdd(&ctr->cell_tiles[first], 1ull);
                return;
            }
        } else if (second != first) {
            const uint32_t cls = probe_class(first, second, ratio);
            const uint32_t k = atomicAdd(&acc->tiles[cls], 1u);
            const bool sampled =
                k < kProbeMinSamples ||
                (splitmix_mix(explore_seed ^ (t * 0xD1B54A32D192ED03ull)) >> 40) < kProbeShare;
            const uint16_t st = (uint16_t)(first | (second << 4) | ((cls % kProbeBuckets) << 8));
            if (!sampled) {
                // Pending: pass 2 commits it.
                probe[t] = st | kProbePending;
                return;
            }
            probe[t] = st | kProbeSampled;
            atomicAdd(&acc->sampled, 1ull);
            // The runner-up runs on the shadow instance: timed, no output.
            enqueue_tile(second, shadow_inst, bi, bj, a, b, queue, qstride, split, bb_ksplit,
                         bb_wide, ctr);
        }
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
    if (defer_bs && split && p == STORM_HIP_CELL_BS)
        *deferred = true;
    else
        enqueue_tile(p, inst, bi, bj, a, b, queue, qstride, split, bb_ksplit, bb_wide, ctr);
    atomicAdd(&ctr->cell_tiles[p], 1ull);
    if (p == STORM_HIP_CELL_BB) {
        // B x B tiles skip the per-pair span test (same cost either way);
        // their pairs are counted here.
        const uint64_t ni = a.rows, nj = b.rows;
        atomicAdd(&ctr->cell_pairs[p], (unsigned long long)(bi == bj ? ni * (ni - 1) / 2 : ni * nj));
    }
    if (fallback) atomicAdd(&ctr->fallback_tiles, 1ull);
}

// One thread per row bi and pair of tile columns (2k, 2k + 1): decides both
// tile pairs, then queues their B x S choices, as wide stripes when both chose
// B x S and bs_wide_ok holds (wide_bs: the call may pair them).
__global__ __launch_bounds__(256) void select_kernel(
    const TileMeta* __restrict__ tm, uint32_t T, DevModel M, uint32_t reps, uint32_t policy,
    uint32_t fixed_cell, const float* __restrict__ learned_cost,
    const uint8_t* __restrict__ learned_ok, uint8_t* __restrict__ decisions,
    QItem* __restrict__ queue, uint64_t qstride, uint32_t inst, uint32_t tile, bool split,
    double thr, uint32_t bb_ksplit, bool bb_wide, uint64_t explore_seed, uint32_t explore_share,
    const uint32_t* __restrict__ sel, DevCounters* __restrict__ ctr, uint16_t* __restrict__ probe,
    ProbeAcc* __restrict__ acc, uint32_t shadow_inst, bool wide_bs,
    const uint64_t* __restrict__ tzone, uint32_t zone_ld) {
    const uint32_t KT = (T + 1) / 2;
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (uint64_t)T * KT) return;
    const uint32_t bi = (uint32_t)(g / KT), bj0 = 2 * (uint32_t)(g % KT);
    if (bj0 + 1 < bi) return;   // both columns below the diagonal
    const bool defer = wide_bs && bi < bj0;
    bool d[2] = {false, false};
    for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t bj = bj0 + h;
        if (bj < bi || bj >= T) continue;
        select_tile(tri_tile_id(bi, bj, T), bi, bj, tm, M, reps, policy, fixed_cell,
                    learned_cost, learned_ok, decisions, queue, qstride, inst, split, thr,
                    bb_ksplit, bb_wide, explore_seed, explore_share, sel, ctr, probe, acc,
                    shadow_inst, defer, &d[h], tzone, zone_ld);
    }
    if (!d[0] && !d[1]) return;
    const TileMeta a = tm[bi];
    if (d[0] && d[1] && bs_wide_ok(bi, bj0, tile, a, tm[bj0], tm[bj0 + 1])) {
        const uint32_t ns = (a.rows + kSplitRows - 1) / kSplitRows;
        uint32_t slot = atomicAdd(&ctr->qcount[STORM_HIP_CELL_BS], ns);
        for (uint32_t x = 0; x < ns; ++x)
            queue[(uint64_t)STORM_HIP_CELL_BS * qstride + slot++] =
                QItem{inst, bi, bj0, make_sub(kSplitRows, x, kWideSide)};
        return;
    }
    for (uint32_t h = 0; h < 2; ++h)
        if (d[h])
            enqueue_tile(STORM_HIP_CELL_BS, inst, bi, bj0 + h, a, tm[bj0 + h], queue, qstride,
                         split, bb_ksplit, bb_wide, ctr);
}

// Probe: add each timed item of pass 1 to its tile's class, as CU-time (wall
// time over the class kernel's workgroups per CU, 10 ns per tick), first or
// second candidate by the queue it ran in. Grid: (x, queue).
__global__ __launch_bounds__(256) void probe_reduce_kernel(
    const DevCorpus* __restrict__ Ds, uint32_t n_inst, const QItem* __restrict__ queue,
    uint64_t qstride, const DevCounters* __restrict__ ctr,
    const unsigned long long* __restrict__ ticks, ProbeBpc bpc, ProbeAcc* __restrict__ acc) {
    const uint32_t q = blockIdx.y;
    const int cell = q == kQueueBBWide ? STORM_HIP_CELL_BB : (int)q;
    if (cell >= 10) return;   // EMPTY, C x B and refine never probe
    const uint32_t count = ctr->qcount[q];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < count;
         k += gridDim.x * blockDim.x) {
        const QItem e = queue[(uint64_t)q * qstride + k];
        const uint32_t inst = e.inst >= n_inst ? e.inst - n_inst : e.inst;
        const DevCorpus& D = Ds[inst];
        const uint32_t T = (D.n + D.tile - 1) / D.tile;
        const uint64_t t = (uint64_t)e.bi * (2ull * T - e.bi + 1) / 2 + (e.bj - e.bi);
        const uint16_t st = D.probe[t];
        if (!(st & kProbeSampled)) continue;
        const int first = st & 15, second = (st >> 4) & 15;
        const uint32_t cls = ((uint32_t)first * 10u + (uint32_t)second) * kProbeBuckets +
                             ((st >> 8) & 15);
        const int which = cell == first ? 0 : (cell == second ? 1 : -1);
        if (which < 0) continue;
        const double ns = (double)ticks[(uint64_t)q * qstride + k] * 10.0 / bpc.v[cell];
        atomicAdd(&acc->ns[which][cls], (unsigned long long)ns);
    }
}

// Pass 2 starts its queues from empty; the counters of pass 1 stay.
__global__ void probe_reset_queues_kernel(DevCounters* ctr) {
    const int q = threadIdx.x;
    if (q < kNumQueues) { ctr->qcount[q] = 0; ctr->qhead[q] = 0; }
}

int launch_probe_reduce(hipStream_t s, const DevCorpus* Ds, uint32_t n_inst, const QItem* queue,
                        uint64_t qstride, const DevCounters* ctr,
                        const unsigned long long* ticks, const ProbeBpc& bpc, ProbeAcc* acc) {
    const uint32_t gx = (uint32_t)std::min<uint64_t>((qstride + 255) / 256, 1024);
    hipLaunchKernelGGL(probe_reduce_kernel, dim3(std::max(gx, 1u), kNumQueues), dim3(256), 0, s,
                       Ds, n_inst, queue, qstride, ctr, ticks, bpc, acc);
    SH_HIP(hipGetLastError());
    hipLaunchKernelGGL(probe_reset_queues_kernel, dim3(1), dim3(64), 0, s,
                       const_cast<DevCounters*>(ctr));
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
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
                  const uint32_t* sel, DevCounters* ctr, uint16_t* probe, ProbeAcc* acc,
                  uint32_t shadow_inst, bool wide_bs, const uint64_t* tzone,
                  uint32_t zone_ld) {
    const uint64_t threads = (uint64_t)T * ((T + 1) / 2);
    if (threads == 0) return STORM_HIP_OK;
    hipLaunchKernelGGL(select_kernel, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, s,
                       tm, T, M, reps, policy, fixed_cell, learned_cost, learned_ok, decisions,
                       queue, qstride, inst, tile, tile > kSplitRows, thr, bb_ksplit, bb_wide,
                       explore_seed, explore_share, sel, ctr, probe, acc, shadow_inst, wide_bs,
                       tzone, zone_ld);
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
