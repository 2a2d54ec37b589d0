// storm::gpu::allpairs_sum / allpairs_tiles over the C ABI: the reference's
// batched driver interface (kernels/storm_allpairs.h:196-214) with every
// pair computed on the GPU engine.
#include "storm_allpairs_gpu.h"

#include <algorithm>
#include <chrono>
#include <vector>

namespace storm {
namespace gpu {

namespace {

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Policy -> options; UNSUPPORTED for the per-pair policies.
int options_for(Policy policy, uint32_t tile, Pairing fixed_cell, storm_hip_options* o) {
    *o = storm_hip_options{};
    // The engine selects per tile of 64, 128 or 256 rows; narrower reference
    // tiles (8, 16, 32) select at 64 (the callback granularity stays `tile`).
    o->tile = (tile == 128 || tile == 256) ? tile : 64;
    o->bb_kernel = STORM_HIP_BB_AUTO;
    switch (policy) {
        case Policy::AllBitmap: o->policy = STORM_HIP_POLICY_ALL_BITMAP; break;
        case Policy::PerTile: o->policy = STORM_HIP_POLICY_ANALYTIC; break;
        case Policy::Fixed:
            o->policy = STORM_HIP_POLICY_FIXED;
            o->fixed_cell = (uint32_t)fixed_cell;
            break;
        case Policy::Probe:
        case Policy::Refine: o->policy = STORM_HIP_POLICY_LEARNED; break;
        default: return STORM_HIP_ERR_UNSUPPORTED;
    }
    return STORM_HIP_OK;
}

// Upload, optionally swap in the caller's model, run. On success *out holds
// the corpus (caller frees).
int run(Engine& eng, const RowsCSR& rows, const storm_hip_cost_model* model, Policy policy,
        uint32_t tile, bool no_zonemap, Pairing fixed_cell, uint32_t out_mode,
        AllPairsStats* st, storm_hip_corpus** out) {
    *out = nullptr;
    if (eng.status() != STORM_HIP_OK) return eng.status();
    storm_hip_options o;
    int rc = options_for(policy, tile, fixed_cell, &o);
    if (rc != STORM_HIP_OK) return rc;
    o.out_mode = out_mode;
    storm_hip_ctx* ctx = eng.ctx();
    storm_hip_cost_model saved{};
    const bool swap = model || no_zonemap;
    if (swap) {
        if ((rc = storm_hip_get_cost_model(ctx, &saved)) != STORM_HIP_OK) return rc;
        storm_hip_cost_model m = model ? *model : saved;
        if (no_zonemap) m.bb_zonemap = 0.0;
        if ((rc = storm_hip_set_cost_model(ctx, &m)) != STORM_HIP_OK) return rc;
    }
    storm_hip_corpus* c = nullptr;
    rc = storm_hip_corpus_from_lists(ctx, rows.offs, rows.pos, rows.n_rows, rows.universe,
                                     STORM_HIP_BITMAP_DENSE | STORM_HIP_BUILD_RUNS |
                                         STORM_HIP_BUILD_EWAH | STORM_HIP_BUILD_COMP,
                                     &c);
    storm_hip_stats s{};
    if (rc == STORM_HIP_OK) rc = storm_hip_allpairs(ctx, c, &o, &s);
    if (swap) (void)storm_hip_set_cost_model(ctx, &saved);
    if (rc != STORM_HIP_OK) {
        if (c) storm_hip_corpus_free(c);
        return rc;
    }
    st->sum = s.sum;
    st->pairs = s.pairs;
    st->decisions = s.decisions;
    st->skipped = s.skipped;
    st->ns_selection = s.ms_select * 1e6;
    for (int k = 0; k < (int)Pairing::COUNT && k < STORM_HIP_NCELL; ++k)
        st->cell_pairs[k] = s.cell_pairs[k];
    *out = c;
    return STORM_HIP_OK;
}

}  // namespace

const char* name_of(Policy p) {
    switch (p) {
        case Policy::AllBitmap: return "all-bitmap";
        case Policy::PerPair: return "per-pair";
        case Policy::PerTile: return "per-tile";
        case Policy::Oracle: return "oracle";
        case Policy::Fixed: return "fixed";
        case Policy::Probe: return "probe";
        case Policy::Refine: return "refine";
    }
    return "?";
}

Engine::Engine(int device) { status_ = storm_hip_ctx_create(device, 0, &ctx_); }

Engine::~Engine() {
    if (ctx_) storm_hip_ctx_free(ctx_);
}

AllPairsStats allpairs_sum(Engine& eng, const RowsCSR& rows, const storm_hip_cost_model* model,
                           Policy policy, uint32_t tile, bool no_zonemap, Pairing fixed_cell) {
    AllPairsStats st;
    if (rows.n_rows < 2) return st;
    const uint64_t t0 = now_ns();
    storm_hip_corpus* c = nullptr;
    st.status = run(eng, rows, model, policy, tile, no_zonemap, fixed_cell, STORM_HIP_OUT_SUM,
                    &st, &c);
    if (c) storm_hip_corpus_free(c);
    st.ns_total = (double)(now_ns() - t0);
    return st;
}

AllPairsStats allpairs_tiles(Engine& eng, const RowsCSR& rows, const storm_hip_cost_model* model,
                             Policy policy, TileCallback cb, void* ctx, uint32_t tile) {
    AllPairsStats st;
    if (rows.n_rows < 2 || !cb) return st;
    if (tile == 0) {
        st.status = STORM_HIP_ERR_INVALID;
        return st;
    }
    const uint64_t t0 = now_ns();
    storm_hip_corpus* c = nullptr;
    st.status = run(eng, rows, model, policy, tile, false, Pairing::BS, STORM_HIP_OUT_UPPER_TRI,
                    &st, &c);
    if (st.status != STORM_HIP_OK) return st;
    const uint64_t n = rows.n_rows;
    std::vector<uint32_t> band, counts((size_t)tile * tile);
    for (uint64_t i0 = 0; i0 < n; i0 += tile) {
        const uint64_t i1 = std::min<uint64_t>(i0 + tile, n);
        // Rows i0 .. i1-1 of the packed triangle are one contiguous range.
        const uint64_t off = i0 * n - i0 * (i0 + 1) / 2;
        const uint64_t end = i1 * n - i1 * (i1 + 1) / 2;
        band.resize(end - off);
        if (end > off) {
            st.status = storm_hip_result_copy_tri(eng.ctx(), c, off, end - off, band.data());
            if (st.status != STORM_HIP_OK) break;
        }
        for (uint64_t j0 = i0; j0 < n; j0 += tile) {
            const uint64_t j1 = std::min<uint64_t>(j0 + tile, n);
            std::fill(counts.begin(), counts.end(), 0u);
            for (uint64_t i = i0; i < i1; ++i) {
                const uint64_t row_at = i * n - i * (i + 1) / 2 - off;   // pair (i, i+1)
                for (uint64_t j = std::max(j0, i + 1); j < j1; ++j)
                    counts[(size_t)(i - i0) * tile + (j - j0)] = band[row_at + (j - i - 1)];
            }
            cb((uint32_t)i0, (uint32_t)j0, (uint32_t)(i1 - i0), (uint32_t)(j1 - j0),
               counts.data(), ctx);
        }
    }
    storm_hip_corpus_free(c);
    st.ns_total = (double)(now_ns() - t0);
    return st;
}

}  // namespace gpu
}  // namespace storm
