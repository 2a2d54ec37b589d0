// The all-pairs call on general (mixed-representation) corpora:
//   select   per-tile metadata -> analytic / learned / fixed decision -> one
//            device queue per cell class            (select.hip)
//   dispatch one persistent launch per cell class draining its queue
//            (cells.hip, bb_valu.hip, bb_mfma.hip)
//   reduce   the u64 sum and the per-cell counters by atomics in the launches
// All counts of work live on the device, so the sequence needs no host round
// trip and can be captured in a hipGraph and replayed.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "general.h"
#include "internal.h"

namespace stormhip {

int launch_tile_meta(hipStream_t s, const DevCorpus& D, uint32_t tile, uint32_t T, TileMeta* out);
int launch_select(hipStream_t s, const TileMeta* tm, uint32_t T, const DevModel& M, uint32_t reps,
                  uint32_t policy, uint32_t fixed_cell, const float* learned_cost,
                  const uint8_t* learned_ok, uint8_t* decisions, uint2* queue, DevCounters* ctr);
int launch_compose_rowmap(hipStream_t s, const uint32_t* order, const uint32_t* bm_row,
                          uint32_t npad, uint32_t* rowmap);
int launch_cell(int cell, hipStream_t s, uint32_t grid, const DevCorpus& D, const DevModel& M,
                const TileMeta* tm, uint32_t tile, const uint2* queue, DevCounters* ctr,
                uint32_t* tri, unsigned long long* sum, unsigned long long* tile_cycles);
int general_materialize(storm_hip_ctx* ctx, storm_hip_corpus* c);
void general_free_arenas(storm_hip_corpus* c);
int learned_scores(storm_hip_ctx* ctx, const TileMeta* tm, uint32_t T, uint64_t P,
                   uint32_t universe, float* cost, uint8_t* ok, bool* have_model);
double host_work_units(int p, const TileMeta& a, const TileMeta& b, double margin);
int launch_features(hipStream_t st, const TileMeta* tm, uint32_t T, uint64_t P, uint32_t universe,
                    float* X);

void default_device_model(DeviceCostModel& m) {
    // default_model (kernels/storm_cost.cpp:482-496) and the default knobs
    // STORM_PROBE_SCALE=0.05, STORM_BR_ORIENT=16, STORM_RR_SCALE=1.
    const double r[10] = {0.116, 0.30, 0.60, 0.55, 0.22, 0.35, 0.40, 0.30, 0.45, 0.50};
    for (int i = 0; i < 10; ++i) m.ns_per_unit[i] = r[i];
    m.ns_fixed = 2.0; m.ns_per_occ_word = 0.30; m.ns_probe_hot = 0; m.ns_probe_cold = 0;
    m.probe_scale = 0.05; m.br_margin = 16.0; m.rr_scale = 1.0; m.bb_zonemap = 1.0;
}

namespace {

struct Work {
    uint32_t tile = 0, T = 0, npad = 0;
    uint64_t P = 0;
    TileMeta* tm = nullptr;
    uint8_t* decisions = nullptr;
    uint2* queue = nullptr;
    DevCounters* ctr = nullptr;
    uint32_t* rowmap = nullptr;
    float* learned_cost = nullptr;
    uint8_t* learned_ok = nullptr;
    unsigned long long* sum = nullptr;
    // Captured select -> dispatch -> reduce graph and what it was built for.
    hipGraphExec_t graph = nullptr;
    uint32_t graph_key[4] = {0, 0, 0, 0};
    uint32_t* graph_tri = nullptr;
};

void free_work(Work* w) {
    if (!w) return;
    if (w->graph) (void)hipGraphExecDestroy(w->graph);
    void* p[] = {w->tm, w->decisions, w->queue, w->ctr, w->rowmap, w->learned_cost,
                 w->learned_ok, w->sum};
    for (void* x : p) if (x) (void)hipFree(x);
    delete w;
}

int ensure_work(storm_hip_corpus* c, uint32_t tile) {
    Work* w = static_cast<Work*>(c->work);
    if (w && w->tile == tile) return STORM_HIP_OK;
    free_work(w);
    c->work = nullptr;
    w = new Work();
    c->work = w;
    w->tile = tile;
    w->T = (c->n_rows + tile - 1) / tile;
    w->P = (uint64_t)w->T * (w->T + 1) / 2;
    w->npad = (c->n_rows + kRowPad - 1) / kRowPad * kRowPad + kRowPad;
    const uint64_t P1 = w->P ? w->P : 1;
    SH_HIP(hipMalloc(&w->tm, sizeof(TileMeta) * (w->T ? w->T : 1)));
    SH_HIP(hipMalloc(&w->decisions, P1));
    SH_HIP(hipMalloc(&w->queue, sizeof(uint2) * P1 * kNumCells));
    SH_HIP(hipMalloc(&w->ctr, sizeof(DevCounters)));
    SH_HIP(hipMalloc(&w->rowmap, sizeof(uint32_t) * w->npad));
    SH_HIP(hipMalloc(&w->learned_cost, sizeof(float) * P1 * 10));
    SH_HIP(hipMalloc(&w->learned_ok, P1));
    SH_HIP(hipMalloc(&w->sum, 8));
    return STORM_HIP_OK;
}

// Zero fill as a kernel (graph replays must not depend on memset nodes).
__global__ void zero_kernel(uint32_t* p, uint32_t n_words) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n_words) p[k] = 0;
}

DevCorpus dev_corpus(const storm_hip_corpus* c, const uint32_t* order) {
    DevCorpus D;
    D.n = c->n_rows; D.universe = c->universe; D.nw = c->n_words; D.ld = c->ld;
    D.list_off = c->d_list_off; D.list = c->d_list;
    D.run_off = c->d_run_off; D.run_start = c->d_run_start; D.run_end = c->d_run_end;
    D.ewah_off = c->d_ewah_off; D.ewah = c->d_ewah;
    D.mk_off = c->d_mk_off; D.mk_word = c->d_mk_word; D.mk_pos = c->d_mk_pos;
    D.comp_off = c->d_comp_off; D.comp = c->d_comp;
    D.bm_row = c->d_bm_row; D.bitmap = c->d_bitmap;
    D.meta = c->d_meta;
    D.perm = order;
    return D;
}

DevModel dev_model(const DeviceCostModel& m) {
    DevModel M;
    for (int i = 0; i < 10; ++i) M.ns_per_unit[i] = m.ns_per_unit[i];
    M.ns_fixed = m.ns_fixed; M.ns_per_occ_word = m.ns_per_occ_word;
    M.ns_probe_hot = m.ns_probe_hot; M.ns_probe_cold = m.ns_probe_cold;
    M.probe_scale = m.probe_scale; M.br_margin = m.br_margin; M.rr_scale = m.rr_scale;
    M.bb_zonemap = m.bb_zonemap;
    return M;
}

// The kernel sequence of one selecting call, on `s`. Timing events (if any)
// bracket selection and the B x B launch.
int enqueue_selecting(storm_hip_ctx* ctx, storm_hip_corpus* c, const storm_hip_options& o,
                      Work* w, uint32_t* tri, bool timed) {
    hipStream_t s = ctx->stream;
    const DevCorpus D = dev_corpus(c, c->d_perm);
    const DevModel M = dev_model(ctx->model);
    const uint32_t reps = (c->has_runs ? 1u : 0u) | (c->has_ewah ? 2u : 0u) | (c->has_comp ? 4u : 0u);
    const uint32_t zero_words = (uint32_t)(sizeof(DevCounters) / 4);
    hipLaunchKernelGGL(zero_kernel, dim3((zero_words + 255) / 256), dim3(256), 0, s,
                       reinterpret_cast<uint32_t*>(w->ctr), zero_words);
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, reinterpret_cast<uint32_t*>(w->sum), 2u);
    SH_HIP(hipGetLastError());
    if (timed) SH_HIP(hipEventRecord(ctx->ev[4], s));
    SH_TRY(launch_tile_meta(s, D, w->tile, w->T, w->tm));
    const float* lc = nullptr;
    const uint8_t* lok = nullptr;
    if (o.policy == STORM_HIP_POLICY_LEARNED) {
        bool have = false;
        SH_TRY(learned_scores(ctx, w->tm, w->T, w->P, c->universe, w->learned_cost, w->learned_ok, &have));
        if (have) { lc = w->learned_cost; lok = w->learned_ok; }
    }
    SH_TRY(launch_select(s, w->tm, w->T, M, reps, o.policy, o.fixed_cell, lc, lok, w->decisions,
                         w->queue, w->ctr));
    SH_TRY(launch_compose_rowmap(s, c->d_perm, c->d_bm_row, w->npad, w->rowmap));
    if (timed) SH_HIP(hipEventRecord(ctx->ev[5], s));
    const uint32_t grid = (uint32_t)std::min<uint64_t>(w->P ? w->P : 1, (uint64_t)ctx->cu_count * 4);
    for (int cell = 0; cell < kNumCells; ++cell) {
        if (cell == STORM_HIP_CELL_BB) continue;
        SH_TRY(launch_cell(cell, s, grid, D, M, w->tm, w->tile, w->queue + (uint64_t)cell * w->P,
                           w->ctr, tri, w->sum, nullptr));
    }
    if (timed) SH_HIP(hipEventRecord(ctx->ev[2], s));
    const uint2* qbb = w->queue + (uint64_t)STORM_HIP_CELL_BB * w->P;
    const uint32_t gbb = (uint32_t)std::min<uint64_t>(w->P ? w->P : 1, (uint64_t)ctx->cu_count * 2);
    if (w->tile == 256)
        SH_TRY(launch_bb_mfma(s, c->d_bitmap, c->ld, c->n_words, w->rowmap, c->d_perm, c->n_rows,
                              qbb, &w->ctr->qcount[STORM_HIP_CELL_BB],
                              &w->ctr->qhead[STORM_HIP_CELL_BB], gbb, tri, w->sum));
    else
        SH_TRY(launch_bb_valu(s, (int)w->tile, c->d_bitmap, c->ld, c->n_words, w->rowmap,
                              c->d_perm, c->n_rows, qbb, &w->ctr->qcount[STORM_HIP_CELL_BB],
                              &w->ctr->qhead[STORM_HIP_CELL_BB], gbb, tri, w->sum));
    if (timed) SH_HIP(hipEventRecord(ctx->ev[3], s));
    return STORM_HIP_OK;
}

}  // namespace

void general_free(storm_hip_corpus* c) {
    free_work(static_cast<Work*>(c->work));
    c->work = nullptr;
    if (!c->dense) general_free_arenas(c);
}

int general_allpairs(storm_hip_ctx* ctx, storm_hip_corpus* c, const storm_hip_options* opt,
                     storm_hip_stats* st, unsigned long long* d_sum) {
    const storm_hip_options& o = *opt;
    const uint32_t tile = o.tile ? o.tile : 64;
    if (tile != 64 && tile != 128 && tile != 256) {
        set_error("tile must be 64, 128 or 256 rows (got %u)", tile);
        return STORM_HIP_ERR_INVALID;
    }
    hipStream_t s = ctx->stream;
    uint32_t* tri = (o.out_mode == STORM_HIP_OUT_UPPER_TRI) ? c->d_tri : nullptr;
    if (c->n_rows < 2) return STORM_HIP_OK;

    if (o.policy == STORM_HIP_POLICY_ALL_BITMAP) {
        // Every pair B x B in identity order, no span test (storm_allpairs.cpp:
        // AllBitmap), bitmaps materialised on the device where missing.
        SH_TRY(general_materialize(ctx, c));
        SH_TRY(ensure_work(c, 256));
        Work* w = static_cast<Work*>(c->work);
        SH_TRY(launch_compose_rowmap(s, c->d_ident, c->d_bm_row, w->npad, w->rowmap));
        SH_HIP(hipEventRecord(ctx->ev[2], s));
        SH_TRY(launch_bb_dense(o.bb_kernel, s, c->d_bitmap, c->ld, c->n_words, w->rowmap,
                               c->n_rows, tri, d_sum, (uint32_t)ctx->cu_count));
        SH_HIP(hipEventRecord(ctx->ev[3], s));
        st->cell_pairs[STORM_HIP_CELL_BB] = st->pairs;
        st->cell_tiles[STORM_HIP_CELL_BB] = w->P;
        st->bb_launches = 1;
        SH_HIP(hipStreamSynchronize(s));
        float ms = 0;
        hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]);
        st->ms_bb = ms;
        return STORM_HIP_OK;
    }
    if (o.policy == STORM_HIP_POLICY_EXHAUSTIVE) {
        set_error("the exhaustive policy is driven through storm_hip_selector_sample");
        return STORM_HIP_ERR_UNSUPPORTED;
    }

    SH_TRY(ensure_work(c, tile));
    Work* w = static_cast<Work*>(c->work);
    if (o.use_graph) {
        const uint32_t key[4] = {o.policy, o.fixed_cell, tile, ctx->model_version};
        if (!w->graph || std::memcmp(key, w->graph_key, sizeof key) != 0 || w->graph_tri != tri) {
            if (w->graph) { (void)hipGraphExecDestroy(w->graph); w->graph = nullptr; }
            hipGraph_t g;
            SH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue_selecting(ctx, c, o, w, tri, false);
            hipError_t e = hipStreamEndCapture(s, &g);
            if (rc != STORM_HIP_OK) return rc;
            if (e != hipSuccess) { set_error("graph capture: %s", hipGetErrorString(e)); return STORM_HIP_ERR_DEVICE; }
            e = hipGraphInstantiate(&w->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) { set_error("graph instantiate: %s", hipGetErrorString(e)); return STORM_HIP_ERR_DEVICE; }
            std::memcpy(w->graph_key, key, sizeof key);
            w->graph_tri = tri;
        }
        SH_HIP(hipGraphLaunch(w->graph, s));
    } else {
        SH_TRY(enqueue_selecting(ctx, c, o, w, tri, true));
    }
    // Fold the path's sum into the caller's accumulator and read the counters.
    SH_HIP(hipMemcpyAsync(d_sum, w->sum, 8, hipMemcpyDeviceToDevice, s));
    DevCounters h;
    SH_HIP(hipMemcpyAsync(&h, w->ctr, sizeof h, hipMemcpyDeviceToHost, s));
    c->last_cells.resize(w->P);
    if (w->P) SH_HIP(hipMemcpyAsync(c->last_cells.data(), w->decisions, w->P, hipMemcpyDeviceToHost, s));
    SH_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < kNumCells; ++k) {
        st->cell_pairs[k] = h.cell_pairs[k];
        st->cell_tiles[k] = h.cell_tiles[k];
    }
    st->skipped = h.skipped;
    st->fallback_tiles = h.fallback_tiles;
    st->decisions = w->P;
    st->bb_launches = h.qcount[STORM_HIP_CELL_BB] ? 1 : 0;
    if (!o.use_graph) {
        float ms = 0;
        hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]);
        st->ms_select = ms;
        hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]);
        st->ms_bb = ms;
    }
    return STORM_HIP_OK;
}

}  // namespace stormhip

using namespace stormhip;

extern "C" {

int storm_hip_set_cost_model(storm_hip_ctx* ctx, const storm_hip_cost_model* m) {
    if (!ctx || !m) return STORM_HIP_ERR_INVALID;
    for (int i = 0; i < 10; ++i) ctx->model.ns_per_unit[i] = m->ns_per_unit[i];
    ctx->model.ns_fixed = m->ns_fixed;
    ctx->model.ns_per_occ_word = m->ns_per_occ_word;
    ctx->model.ns_probe_hot = m->ns_probe_hot;
    ctx->model.ns_probe_cold = m->ns_probe_cold;
    ctx->model.bb_zonemap = m->bb_zonemap;
    ++ctx->model_version;
    return STORM_HIP_OK;
}

int storm_hip_get_cost_model(storm_hip_ctx* ctx, storm_hip_cost_model* m) {
    if (!ctx || !m) return STORM_HIP_ERR_INVALID;
    for (int i = 0; i < 10; ++i) m->ns_per_unit[i] = ctx->model.ns_per_unit[i];
    m->ns_fixed = ctx->model.ns_fixed;
    m->ns_per_occ_word = ctx->model.ns_per_occ_word;
    m->ns_probe_hot = ctx->model.ns_probe_hot;
    m->ns_probe_cold = ctx->model.ns_probe_cold;
    m->bb_zonemap = ctx->model.bb_zonemap;
    return STORM_HIP_OK;
}

int storm_hip_selector_sample(storm_hip_ctx* ctx, storm_hip_corpus* c, uint32_t tile,
                              uint32_t max_tiles, uint64_t seed, uint32_t repeats,
                              float* features, float* log2_ns, uint8_t* mask, uint32_t* tile_index,
                              uint32_t* n_sampled) {
    if (!ctx || !c || c->ctx != ctx || c->dense || !n_sampled) return STORM_HIP_ERR_INVALID;
    if (max_tiles && (!features || !log2_ns || !mask || !tile_index)) return STORM_HIP_ERR_NULL_VALUES;
    if (tile != 64 && tile != 128 && tile != 256) return STORM_HIP_ERR_INVALID;
    *n_sampled = 0;
    if (c->n_rows < 2 || max_tiles == 0) return STORM_HIP_OK;
    SH_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    SH_TRY(ensure_work(c, tile));
    Work* w = static_cast<Work*>(c->work);
    const int F = STORM_HIP_MLP_FEATURES;
    // Tile metas, features and the analytic pass (for the proven-empty tiles).
    storm_hip_options o{};
    o.policy = STORM_HIP_POLICY_ANALYTIC;
    o.tile = tile;
    const DevCorpus D = dev_corpus(c, c->d_perm);
    const DevModel M = dev_model(ctx->model);
    const uint32_t reps = (c->has_runs ? 1u : 0u) | (c->has_ewah ? 2u : 0u) | (c->has_comp ? 4u : 0u);
    SH_HIP(hipMemsetAsync(w->ctr, 0, sizeof(DevCounters), s));
    SH_TRY(launch_tile_meta(s, D, w->tile, w->T, w->tm));
    SH_TRY(launch_select(s, w->tm, w->T, M, reps, o.policy, 0, nullptr, nullptr, w->decisions,
                         w->queue, w->ctr));
    SH_TRY(launch_compose_rowmap(s, c->d_perm, c->d_bm_row, w->npad, w->rowmap));
    float* dX = nullptr;
    SH_HIP(hipMalloc(&dX, sizeof(float) * F * w->P));
    std::vector<float> X((size_t)F * w->P);
    std::vector<TileMeta> tm(w->T);
    std::vector<uint8_t> dec(w->P);
    int rc = launch_features(s, w->tm, w->T, w->P, c->universe, dX);
    if (rc == STORM_HIP_OK) {
        (void)hipMemcpyAsync(X.data(), dX, X.size() * 4, hipMemcpyDeviceToHost, s);
        (void)hipMemcpyAsync(tm.data(), w->tm, tm.size() * sizeof(TileMeta), hipMemcpyDeviceToHost, s);
        (void)hipMemcpyAsync(dec.data(), w->decisions, w->P, hipMemcpyDeviceToHost, s);
        if (hipStreamSynchronize(s) != hipSuccess) rc = STORM_HIP_ERR_DEVICE;
    }
    (void)hipFree(dX);
    if (rc != STORM_HIP_OK) return rc;
    // Seeded sample of the non-empty tile pairs.
    std::vector<uint32_t> cand;
    for (uint64_t t = 0; t < w->P; ++t)
        if (dec[t] != STORM_HIP_CELL_EMPTY) cand.push_back((uint32_t)t);
    uint64_t rs = seed ? seed : kGolden;
    for (size_t i = cand.size(); i > 1; --i) {
        rs += kGolden;
        std::swap(cand[i - 1], cand[splitmix_mix(rs) % i]);
    }
    if (cand.size() > max_tiles) cand.resize(max_tiles);
    std::sort(cand.begin(), cand.end());
    const uint32_t S = (uint32_t)cand.size();
    std::vector<uint2> pairs(S);
    std::vector<uint32_t> elig(S);
    for (uint32_t k = 0; k < S; ++k) {
        uint32_t bi, bj;
        tri_tile(cand[k], w->T, bi, bj);
        pairs[k] = make_uint2(bi, bj);
        elig[k] = eligible_mask(tm[bi], tm[bj], reps);
        tile_index[k] = cand[k];
        std::memcpy(features + (size_t)k * F, &X[(size_t)cand[k] * F], F * 4);
    }
    // Time every eligible cell on its sampled tiles, each launch alone.
    unsigned long long* dticks = nullptr;
    SH_HIP(hipMalloc(&dticks, sizeof(unsigned long long) * std::max<uint32_t>(S, 1)));
    std::vector<unsigned long long> ticks(S);
    std::vector<double> best((size_t)S * 10, 0.0);
    std::vector<uint8_t> has((size_t)S * 10, 0);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(std::max<uint32_t>(S, 1), (uint64_t)ctx->cu_count * 4);
    for (int cell = 0; cell < 10 && rc == STORM_HIP_OK; ++cell) {
        std::vector<uint2> q;
        std::vector<uint32_t> slot_of;
        for (uint32_t k = 0; k < S; ++k)
            if (elig[k] >> cell & 1) { q.push_back(pairs[k]); slot_of.push_back(k); }
        if (q.empty()) continue;
        uint2* dq = w->queue + (uint64_t)cell * w->P;
        (void)hipMemcpyAsync(dq, q.data(), q.size() * sizeof(uint2), hipMemcpyHostToDevice, s);
        for (uint32_t r = 0; r < std::max(1u, repeats) && rc == STORM_HIP_OK; ++r) {
            DevCounters h{};
            h.qcount[cell] = (uint32_t)q.size();
            (void)hipMemcpyAsync(w->ctr, &h, sizeof h, hipMemcpyHostToDevice, s);
            if (cell == STORM_HIP_CELL_BB) {
                if (w->tile == 256)
                    rc = launch_bb_mfma(s, c->d_bitmap, c->ld, c->n_words, w->rowmap, c->d_perm,
                                        c->n_rows, dq, &w->ctr->qcount[cell], &w->ctr->qhead[cell],
                                        grid, nullptr, w->sum, dticks);
                else
                    rc = launch_bb_valu(s, (int)w->tile, c->d_bitmap, c->ld, c->n_words, w->rowmap,
                                        c->d_perm, c->n_rows, dq, &w->ctr->qcount[cell],
                                        &w->ctr->qhead[cell], grid, nullptr, w->sum, dticks);
            } else {
                rc = launch_cell(cell, s, grid, D, M, w->tm, w->tile, dq, w->ctr, nullptr, w->sum,
                                 dticks);
            }
            if (rc != STORM_HIP_OK) break;
            (void)hipMemcpyAsync(ticks.data(), dticks, q.size() * 8, hipMemcpyDeviceToHost, s);
            if (hipStreamSynchronize(s) != hipSuccess) { rc = STORM_HIP_ERR_DEVICE; break; }
            for (size_t e = 0; e < q.size(); ++e) {
                const uint32_t k = slot_of[e];
                const double ns = (double)ticks[e] * 10.0;   // 100 MHz wall clock
                if (!has[(size_t)k * 10 + cell] || ns < best[(size_t)k * 10 + cell])
                    best[(size_t)k * 10 + cell] = ns;
                has[(size_t)k * 10 + cell] = 1;
            }
        }
    }
    (void)hipFree(dticks);
    if (rc != STORM_HIP_OK) return rc;
    for (uint32_t k = 0; k < S; ++k)
        for (int cell = 0; cell < 10; ++cell) {
            const size_t i = (size_t)k * 10 + cell;
            mask[i] = has[i];
            log2_ns[i] = has[i] ? (float)std::log2(std::max(best[i], 10.0)) : 0.f;
        }
    *n_sampled = S;
    return STORM_HIP_OK;
}

int storm_hip_calibrate(storm_hip_ctx* ctx, storm_hip_corpus* c, uint32_t tile,
                        uint32_t max_tiles, uint64_t seed, uint32_t repeats,
                        storm_hip_cost_model* out) {
    if (!ctx || !c || c->dense) return STORM_HIP_ERR_INVALID;
    const int F = STORM_HIP_MLP_FEATURES;
    std::vector<float> X((size_t)max_tiles * F), Y((size_t)max_tiles * 10);
    std::vector<uint8_t> Mk((size_t)max_tiles * 10);
    std::vector<uint32_t> idx(max_tiles);
    uint32_t n = 0;
    SH_TRY(storm_hip_selector_sample(ctx, c, tile, max_tiles, seed, repeats, X.data(), Y.data(),
                                     Mk.data(), idx.data(), &n));
    Work* w = static_cast<Work*>(c->work);
    if (!w) return STORM_HIP_ERR_INVALID;
    std::vector<TileMeta> tm(w->T);
    SH_HIP(hipMemcpy(tm.data(), w->tm, tm.size() * sizeof(TileMeta), hipMemcpyDeviceToHost));
    DeviceCostModel m = ctx->model;
    for (int p = 0; p < 10; ++p) {
        std::vector<double> r;
        for (uint32_t k = 0; k < n; ++k) {
            if (!Mk[(size_t)k * 10 + p]) continue;
            uint32_t bi, bj;
            tri_tile(idx[k], w->T, bi, bj);
            const double na = tm[bi].rows, nb = tm[bj].rows;
            const double pairs = bi == bj ? na * (na - 1) / 2 : na * nb;
            const double units = std::max(1.0, host_work_units(p, tm[bi], tm[bj], m.br_margin));
            if (pairs < 1) continue;
            r.push_back(std::exp2((double)Y[(size_t)k * 10 + p]) / (pairs * units));
        }
        if (r.empty()) { m.ns_per_unit[p] = 0.0; continue; }
        std::nth_element(r.begin(), r.begin() + r.size() / 2, r.end());
        m.ns_per_unit[p] = r[r.size() / 2];
    }
    m.ns_fixed = 0.0; m.ns_per_occ_word = 0.0; m.ns_probe_hot = 0.0; m.ns_probe_cold = 0.0;
    m.bb_zonemap = 0.0;
    ctx->model = m;
    ++ctx->model_version;
    if (out) SH_TRY(storm_hip_get_cost_model(ctx, out));
    return STORM_HIP_OK;
}

int storm_hip_last_decisions(storm_hip_ctx* ctx, const storm_hip_corpus* c, uint8_t* cells,
                             uint64_t n_tiles) {
    if (!ctx || !c) return STORM_HIP_ERR_INVALID;
    if (!cells && n_tiles) return STORM_HIP_ERR_NULL_VALUES;
    if (n_tiles > c->last_cells.size()) return STORM_HIP_ERR_RANGE;
    std::memcpy(cells, c->last_cells.data(), n_tiles);
    return STORM_HIP_OK;
}

}  // extern "C"
