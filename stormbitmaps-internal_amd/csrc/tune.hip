// Autotuner: times candidate launch shapes on the caller's own corpus and
// keeps the fastest. Shapes never change counts, only speed.
#include <cstring>
#include <vector>

#include "internal.h"

using namespace stormhip;

namespace {

bool valid(const storm_hip_bb_config& c) {
    return (c.stage_words == 4 || c.stage_words == 8 || (c.stage_words == 16 && c.waves == 8)) &&
           c.band <= 64 && c.blocks_per_cu >= 1 && c.blocks_per_cu <= 16 &&
           (c.waves == 4 || c.waves == 8);
}

MfmaConfig to_internal(const storm_hip_bb_config& c) {
    MfmaConfig m;
    m.stage_words = c.stage_words;
    m.band = c.band;
    m.blocks_per_cu = c.blocks_per_cu;
    m.waves = c.waves;
    return m;
}

storm_hip_bb_config to_public(const MfmaConfig& m) {
    storm_hip_bb_config c{};
    c.stage_words = m.stage_words;
    c.band = m.band;
    c.blocks_per_cu = m.blocks_per_cu;
    c.waves = m.waves;
    return c;
}

bool same(const MfmaConfig& a, const MfmaConfig& b) {
    return a.stage_words == b.stage_words && a.band == b.band &&
           a.blocks_per_cu == b.blocks_per_cu && a.waves == b.waves;
}

// One call with sums only; its device time in ms.
int time_call(storm_hip_ctx* ctx, storm_hip_corpus* c, const storm_hip_options& o, double* ms) {
    storm_hip_stats st{};
    SH_TRY(storm_hip_allpairs(ctx, c, &o, &st));
    *ms = st.ms_device;
    return STORM_HIP_OK;
}

}  // namespace

extern "C" {

int storm_hip_set_bb_config(storm_hip_ctx* ctx, const storm_hip_bb_config* cfg) {
    if (!ctx || !cfg || !valid(*cfg)) {
        set_error("bb config: stage_words 4, 8 or 16 (16: 8 waves), band <= 64, "
                  "blocks_per_cu in [1, 16], waves 4 or 8");
        return STORM_HIP_ERR_INVALID;
    }
    ctx->mfma = to_internal(*cfg);
    ++ctx->model_version;   // captured graphs bake the launch shape in
    return STORM_HIP_OK;
}

int storm_hip_get_bb_config(storm_hip_ctx* ctx, storm_hip_bb_config* cfg) {
    if (!ctx || !cfg) return STORM_HIP_ERR_INVALID;
    *cfg = to_public(ctx->mfma);
    return STORM_HIP_OK;
}

int storm_hip_autotune(storm_hip_ctx* ctx, storm_hip_corpus* c, const storm_hip_options* opt,
                       storm_hip_tune_result* result) {
    if (!ctx || !c || c->ctx != ctx || !result) return STORM_HIP_ERR_INVALID;
    std::memset(result, 0, sizeof *result);
    storm_hip_options o{};
    if (opt) o = *opt;
    o.out_mode = STORM_HIP_OUT_SUM;
    if (o.bb_kernel == STORM_HIP_BB_VALU && c->dense) {
        // The VALU form has no tunable shape.
        result->bb = to_public(ctx->mfma);
        double ms = 0;
        SH_TRY(time_call(ctx, c, o, &ms));
        SH_TRY(time_call(ctx, c, o, &ms));
        result->ms_best = result->ms_default = ms;
        result->n_tried = 1;
        return STORM_HIP_OK;
    }
    const MfmaConfig start = ctx->mfma;
    const uint32_t start_tile = o.tile ? o.tile : 64;
    int rc = STORM_HIP_OK;
    double best = 0;
    uint32_t tile = 0;
    // Selection tile first (general corpora), with the starting shape.
    if (!c->dense && o.policy != STORM_HIP_POLICY_ALL_BITMAP) {
        for (uint32_t t : {64u, 128u, 256u}) {
            o.tile = t;
            double ms = 0;
            if ((rc = time_call(ctx, c, o, &ms)) != STORM_HIP_OK) return rc;   // warm-up
            if ((rc = time_call(ctx, c, o, &ms)) != STORM_HIP_OK) return rc;
            ++result->n_tried;
            if (t == start_tile) result->ms_default = ms;
            if (!tile || ms < best) { best = ms; tile = t; }
        }
        o.tile = tile;
        result->tile = tile;
        if (tile != 256) {
            result->bb = to_public(start);
            result->ms_best = best;
            return STORM_HIP_OK;
        }
    } else {
        o.tile = start_tile;
    }
    // Matrix-core B x B shape. The K-split target only matters for corpora
    // with fewer than a few 256-row tiles per CU; elsewhere it is not varied.
    const uint64_t T = (c->n_rows + 255) / 256, tiles = T * (T + 1) / 2;
    std::vector<uint32_t> bpcs = {start.blocks_per_cu};
    if (tiles < 4ull * (uint64_t)ctx->cu_count) bpcs = {1u, 2u, 4u};
    MfmaConfig best_cfg = start;
    bool have = false;
    std::vector<MfmaConfig> cands;
    for (uint32_t waves : {8u, 4u})
        for (uint32_t sw : {16u, 8u, 4u})
            for (uint32_t band : {0u, 2u, 4u, 8u})
                for (uint32_t bpc : bpcs) {
                    if (sw == 16 && waves != 8) continue;   // 16-word stages: 8 waves only
                    MfmaConfig m;
                    m.stage_words = sw;
                    m.band = band;
                    m.blocks_per_cu = bpc;
                    m.waves = waves;
                    cands.push_back(m);
                }
    for (const MfmaConfig& m : cands) {
        ctx->mfma = m;
        ++ctx->model_version;
        double ms = 0;
        rc = time_call(ctx, c, o, &ms);
        if (rc == STORM_HIP_OK) rc = time_call(ctx, c, o, &ms);
        if (rc != STORM_HIP_OK) break;
        ++result->n_tried;
        if (same(m, start) && c->dense) result->ms_default = ms;
        if (!have || ms < best) { best = ms; best_cfg = m; have = true; }
    }
    ctx->mfma = rc == STORM_HIP_OK ? best_cfg : start;
    ++ctx->model_version;
    if (rc != STORM_HIP_OK) return rc;
    result->bb = to_public(best_cfg);
    result->ms_best = best;
    return STORM_HIP_OK;
}

}  // extern "C"
