// The analytic cost model of kernels/storm_cost.cpp ported to the device:
// work units per cell and the predicted cost of a cell on a pair of tiles
// (or of rows, as one-row tiles). Shared by the selection and cell kernels.
#pragma once

#include "general.h"

namespace stormhip {

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

__host__ __device__ inline double work_units(int p, const TileMeta& a, const TileMeta& b, double margin) {
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

__host__ __device__ inline double predict(const DevModel& m, int p, const TileMeta& a, const TileMeta& b) {
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
__device__ inline int select_pairing(const DevModel& m, const TileMeta& a, const TileMeta& b, double* cost) {
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

}  // namespace stormhip
