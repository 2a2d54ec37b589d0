// The all-pairs call on general (mixed-representation) corpora, for one corpus
// or a batch of independent corpora in the same launches:
//   select   per corpus: tile metadata -> analytic / learned / fixed decision
//            per tile pair -> appended to ONE device queue per cell class,
//            entries tagged with their corpus             (select.hip)
//   dispatch one persistent launch per cell class drains its queue across all
//            corpora of the call (cells.hip, bb_valu.hip, bb_mfma.hip)
//   reduce   per-corpus sums, per-cell counters and per-corpus finish times
//            by atomics inside those launches
// Every count of work lives on the device, so the whole sequence is captured
// once in a hipGraph and replayed.
#include <algorithm>
#include <chrono>
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
                  const uint8_t* learned_ok, uint8_t* decisions, QItem* queue, uint64_t qstride,
                  uint32_t inst, uint32_t tile, double thr, uint32_t bb_ksplit,
                  bool bb_wide, uint64_t explore_seed, uint32_t explore_share,
                  const uint32_t* sel, DevCounters* ctr, uint16_t* probe, ProbeAcc* acc,
                  uint32_t shadow_inst, bool wide_bs, const uint64_t* tzone,
                  uint32_t zone_ld);
int launch_refine(hipStream_t s, uint32_t grid, const DevCorpus* Ds, const DevModel& M,
                  const QItem* queue, DevCounters* ctr, unsigned long long* tile_cycles);
uint32_t refine_blocks_per_cu();
int launch_compose_rowmap(hipStream_t s, const uint32_t* order, const uint32_t* bm_row,
                          uint32_t npad, uint32_t* rowmap);
int launch_cell(int cell, hipStream_t s, uint32_t grid, const DevCorpus* Ds, const DevModel& M,
                const QItem* queue, DevCounters* ctr, unsigned long long* tile_cycles);
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

// Per-corpus selection state for one tile width.
struct Work {
    uint32_t tile = 0, T = 0, npad = 0;
    uint64_t P = 0;
    TileMeta* tm = nullptr;
    uint8_t* decisions = nullptr;
    uint32_t* rowmap = nullptr;
    float* learned_cost = nullptr;
    uint8_t* learned_ok = nullptr;
    uint16_t* probe = nullptr;        // per-tile probe state (PROBE calls)
    uint64_t* tzone = nullptr;        // per row tile: OR of its rows' zone maps
};

void free_work(Work* w) {
    if (!w) return;
    void* p[] = {w->tm, w->decisions, w->rowmap, w->learned_cost, w->learned_ok, w->probe,
                 w->tzone};
    for (void* x : p) if (x) (void)hipFree(x);
    delete w;
}

// Selection state of corpus c at this tile width (owned by the caller).
int make_work(const storm_hip_corpus* c, uint32_t tile, Work** out) {
    Work* w = new Work();
    *out = w;
    w->tile = tile;
    w->T = (c->n_rows + tile - 1) / tile;
    w->P = (uint64_t)w->T * (w->T + 1) / 2;
    w->npad = (c->n_rows + kRowPad - 1) / kRowPad * kRowPad + kRowPad;
    const uint64_t P1 = w->P ? w->P : 1;
    SH_HIP(hipMalloc(&w->tm, sizeof(TileMeta) * (w->T ? w->T : 1)));
    SH_HIP(hipMalloc(&w->decisions, P1));
    SH_HIP(hipMalloc(&w->rowmap, sizeof(uint32_t) * w->npad));
    SH_HIP(hipMalloc(&w->learned_cost, sizeof(float) * P1 * 10));
    SH_HIP(hipMalloc(&w->learned_ok, P1));
    SH_HIP(hipMalloc(&w->probe, sizeof(uint16_t) * P1));
    if (c->d_zone)
        SH_HIP(hipMalloc(&w->tzone, sizeof(uint64_t) * c->zone_ld * (w->T ? w->T : 1)));
    return STORM_HIP_OK;
}

// Output of a thresholded single-corpus call: the Jaccard threshold and the
// device pair buffer with its counter.
struct Thresh {
    double thr = 0.0;
    PairHit* hits = nullptr;
    unsigned long long* n_hits = nullptr;
    uint64_t cap = 0;
    // Band / cross output (kind 1 / 2, see DevCorpus) and the call's order.
    uint32_t kind = 0, cols = 0, split = 0;
    uint32_t* buf = nullptr;
    uint64_t buf_bytes = 0;
    const uint32_t* inv = nullptr;
    bool identity = false;    // identity order instead of the density order
    // Unthresholded calls: the triangle is packed over density positions.
    bool tri_density = false;
    bool emits() const { return hits != nullptr || kind != 0; }
};

// Device state of one call over a list of corpora: per-corpus selection
// state, the device copy of every corpus's view, the queues and the results.
// Cached and reused (graph included) while the corpora, their device arrays,
// the tile width and the outputs stay the same.
struct Plan {
    Thresh th;                              // thresholded call (th.hits set) or not
    uint32_t* thr_buf = nullptr;            // thresholded 256-row B x B: block buffers
    uint32_t thr_grid = 0;                  // blocks (one 256 x 256 buffer each)
    std::vector<storm_hip_corpus*> cs;
    std::vector<uint32_t*> tris;
    std::vector<Work*> works;               // one per corpus, owned
    std::vector<DevCorpus> h_corpora;       // host copy of d_corpora
    uint32_t tile = 0;
    uint64_t totalP = 0;
    uint64_t qstride = 0;                   // queue slots per cell class
    DevCorpus* d_corpora = nullptr;
    QItem* queue = nullptr;                 // kNumQueues x qstride
    DevCounters* ctr = nullptr;
    unsigned long long* sums = nullptr;     // one per corpus
    unsigned long long* finish = nullptr;   // one per corpus, then the start stamp
    hipGraphExec_t graph = nullptr;
    uint32_t graph_key[4] = {0, 0, 0, 0};
    // Exploration calls (eager only): the seed of the cell draws and the
    // per-item times, kNumQueues x qstride, indexed like the queues.
    uint64_t explore_seed = 0;
    uint32_t explore_share = 0;            // tiles exploring, out of 2^24
    // Tile pairs (bi, bj) queued: bi in [sel[0], sel[1]), bj in [sel[2], sel[3]).
    // The select kernel reads d_sel, so one captured graph serves every range;
    // sel_dev mirrors what d_sel holds.
    uint32_t sel[4] = {0, 0xFFFFFFFFu, 0, 0xFFFFFFFFu};
    uint32_t sel_dev[4] = {0, 0, 0, 0};
    uint32_t* d_sel = nullptr;
    unsigned long long* ticks = nullptr;
    // Probe calls: per-class measurements, and the shadow instances' sum
    // (d_corpora holds each corpus's shadow at index k + n).
    ProbeAcc* acc = nullptr;
    unsigned long long* shadow_sum = nullptr;
};

void free_plan(Plan* p) {
    if (!p) return;
    if (p->graph) (void)hipGraphExecDestroy(p->graph);
    for (Work* w : p->works) free_work(w);
    void* q[] = {p->d_corpora, p->queue,   p->ctr,   p->sums,  p->finish,
                 p->thr_buf,   p->ticks,   p->d_sel, p->acc,   p->shadow_sum};
    for (void* x : q) if (x) (void)hipFree(x);
    delete p;
}

uint32_t reps_of(const storm_hip_corpus* c);

DevCorpus dev_corpus(const storm_hip_corpus* c, const Work* w, uint32_t* tri,
                     unsigned long long* sum, unsigned long long* finish, const Thresh& th) {
    DevCorpus D{};
    D.thr = th.thr; D.hits = th.hits; D.n_hits = th.n_hits; D.hit_cap = th.cap;
    D.n = c->n_rows; D.universe = c->universe; D.nw = c->n_words; D.ld = c->ld;
    D.list_off = c->d_list_off; D.list = c->d_list;
    D.run_off = c->d_run_off; D.run_start = c->d_run_start; D.run_end = c->d_run_end;
    D.ewah_off = c->d_ewah_off; D.ewah = c->d_ewah;
    D.mk_off = c->d_mk_off; D.mk_word = c->d_mk_word; D.mk_pos = c->d_mk_pos;
    D.comp_off = c->d_comp_off; D.comp = c->d_comp;
    D.bm_row = c->d_bm_row; D.bitmap = c->d_bitmap;
    D.rank_row = c->d_rank_row; D.rank = c->d_rank; D.rank_ld = c->rank_ld;
    D.occ_row = c->d_occ_row; D.occ = c->d_occ; D.occ_ld = c->occ_ld;
    D.meta = c->d_meta;
    D.perm = th.identity ? c->d_ident : c->d_perm;
    D.pspan = th.identity ? nullptr : reinterpret_cast<const uint4*>(c->d_pspan);
    D.emit_kind = th.kind; D.emit_cols = th.cols; D.emit_split = th.split;
    D.emit_buf = th.buf; D.inv = th.inv;
    D.slice = c->d_slice; D.slice_of = c->d_slice_of;
    D.zone = c->d_zone; D.zone_ld = c->zone_ld; D.zone_shift = c->zone_shift;
    D.tzone = w->tzone;
    D.tm = w->tm; D.rowmap = w->rowmap; D.tile = w->tile; D.probe = w->probe;
    D.reps = reps_of(c);
    D.tri = tri; D.sum = sum; D.finish = finish;
    if (tri && th.tri_density) { D.tri_pos = 1; D.inv = c->d_inv; }
    // K slices for wide 256-row B x B tiles (not in thresholded calls, whose
    // B x B tiles test each pair from the whole count).
    D.bb_ksplit = (w->tile == 256 && !th.emits()) ? bb_slices_for(c->n_words) : 1;
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

uint32_t reps_of(const storm_hip_corpus* c) {
    return (c->has_runs ? kRepRuns : 0u) | (c->has_ewah ? kRepEwah : 0u) |
           (c->has_comp ? kRepComp : 0u) | (c->has_lists ? kRepLists : 0u);
}

// (Re)build the plan for this list of corpora, tile width and outputs.
// A plan is rebuilt when anything it baked in moved: the corpus list, the
// outputs, the tile width, or any device array of a corpus (materialising
// bitmaps replaces the bitmap store, for instance).
bool plan_current(const Plan* p, const std::vector<storm_hip_corpus*>& cs,
                  const std::vector<uint32_t*>& tris, uint32_t tile, const Thresh& th) {
    if (!p || p->cs != cs || p->tris != tris || p->tile != tile) return false;
    for (size_t k = 0; k < cs.size(); ++k) {
        const DevCorpus D =
            dev_corpus(cs[k], p->works[k], tris[k], p->sums + k, p->finish + k, th);
        if (std::memcmp(&D, &p->h_corpora[k], sizeof D) != 0) return false;
    }
    return true;
}

int build_plan(Plan*& p, const std::vector<storm_hip_corpus*>& cs,
               const std::vector<uint32_t*>& tris, uint32_t tile, const Thresh& th) {
    free_plan(p);
    p = new Plan();
    p->th = th;
    p->cs = cs;
    p->tris = tris;
    p->tile = tile;
    const size_t n = cs.size();
    for (storm_hip_corpus* c : cs) {
        Work* w = nullptr;
        const int rc = make_work(c, tile, &w);
        p->works.push_back(w);
        SH_TRY(rc);
        p->totalP += w->P;
    }
    const uint64_t P1 = p->totalP ? p->totalP : 1;
    // A tile of a per-pair cell may be queued as (tile / kSplitRows)^2 blocks,
    // a B x B tile as up to kBBMaxSlices K slices.
    uint64_t split = tile > kSplitRows ? (uint64_t)(tile / kSplitRows) * (tile / kSplitRows) : 1;
    for (storm_hip_corpus* c : cs)
        if (tile == 256 && !th.emits()) split = std::max<uint64_t>(split, bb_slices_for(c->n_words));
    p->qstride = P1 * split;
    SH_HIP(hipMalloc(&p->d_corpora, sizeof(DevCorpus) * std::max<size_t>(2 * n, 1)));
    SH_HIP(hipMalloc(&p->queue, sizeof(QItem) * p->qstride * kNumQueues));
    SH_HIP(hipMalloc(&p->ctr, sizeof(DevCounters)));
    SH_HIP(hipMalloc(&p->d_sel, sizeof p->sel));
    SH_HIP(hipMalloc(&p->acc, sizeof(ProbeAcc)));
    SH_HIP(hipMalloc(&p->shadow_sum, 8));
    SH_HIP(hipMalloc(&p->sums, 8 * std::max<size_t>(n, 1)));
    SH_HIP(hipMalloc(&p->finish, 8 * (n + 1)));
    p->h_corpora.resize(n);
    for (size_t k = 0; k < n; ++k)
        p->h_corpora[k] =
            dev_corpus(cs[k], p->works[k], tris[k], p->sums + k, p->finish + k, th);
    // Shadow instances: the same arrays, no output of any kind.
    std::vector<DevCorpus> all(p->h_corpora);
    for (size_t k = 0; k < n; ++k) {
        DevCorpus D = p->h_corpora[k];
        D.tri = nullptr;
        D.hits = nullptr; D.n_hits = nullptr; D.hit_cap = 0;
        D.emit_kind = 0; D.emit_buf = nullptr;
        D.sum = p->shadow_sum;
        D.finish = nullptr;
        D.shadow = 1;
        all.push_back(D);
    }
    if (n)
        SH_HIP(hipMemcpy(p->d_corpora, all.data(), sizeof(DevCorpus) * 2 * n,
                         hipMemcpyHostToDevice));
    return STORM_HIP_OK;
}

// On failure the plan is freed and left null (never half built).
int ensure_plan(Plan*& p, const std::vector<storm_hip_corpus*>& cs,
                const std::vector<uint32_t*>& tris, uint32_t tile, const Thresh& th = Thresh{}) {
    if (plan_current(p, cs, tris, tile, th)) return STORM_HIP_OK;
    const int rc = build_plan(p, cs, tris, tile, th);
    if (rc != STORM_HIP_OK) { free_plan(p); p = nullptr; }
    return rc;
}

// The all-bitmap path's row map (identity order through the bitmap store).
int ensure_bitmap_rowmap(storm_hip_corpus* c, Work** out) {
    Work* w = static_cast<Work*>(c->work);
    if (!w) {
        const int rc = make_work(c, 256, &w);
        if (rc != STORM_HIP_OK) { free_work(w); return rc; }
        c->work = w;
    }
    *out = w;
    return STORM_HIP_OK;
}

// Zero fill as a kernel (graph replays must not depend on memset nodes).
__global__ void zero_kernel(uint32_t* p, uint64_t n_words) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n_words;
         k += (uint64_t)gridDim.x * blockDim.x)
        p[k] = 0;
}
__global__ void stamp_kernel(unsigned long long* out) { *out = wall_clock64(); }

int zero_words(hipStream_t s, void* p, uint64_t bytes) {
    const uint64_t n = bytes / 4;
    if (!n) return STORM_HIP_OK;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 1u << 16);
    hipLaunchKernelGGL(zero_kernel, dim3(grid), dim3(256), 0, s, reinterpret_cast<uint32_t*>(p), n);
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

// Calls whose pairs all go through emit_pair (thresholded, band, cross).
bool thresholded_plan(const Plan* p) { return p->th.emits(); }

// A corpus whose 256-row B x B tiles go to the int8 queue: counts can reach
// 2^24 (2^18 words or more). Thresholded calls and explicit int8 use one
// queue for every corpus.
bool bb_wide(const storm_hip_corpus* c, const storm_hip_options& o, uint32_t tile, bool thr) {
    return tile == 256 && !thr && o.bb_kernel != STORM_HIP_BB_MFMA && c->n_words >= (1u << 18) &&
           bb_item_words(c->n_words, bb_slices_for(c->n_words)) >= (1u << 18);
}

// B x S tile pairs may be queued as wide stripes (two bitmap tiles per
// stripe): 256-row tiles of a corpus with slices, under a policy that decides
// every tile in one pass (not Probe, Refine or PerPair).
bool bs_wide_call(const storm_hip_corpus* c, uint32_t policy, uint32_t tile) {
    return tile == 256 && c->d_slice != nullptr &&
           (policy == STORM_HIP_POLICY_ANALYTIC || policy == STORM_HIP_POLICY_LEARNED ||
            policy == STORM_HIP_POLICY_FIXED || policy == kPolicyExplore);
}

// Dispatch the queued work items: every queue's persistent kernel, on three
// streams, joined back into the context's stream.
int dispatch_queues(storm_hip_ctx* ctx, Plan* p, const storm_hip_options& o, const DevModel& M,
                    bool fp4, bool timed, bool with_ticks) {
    hipStream_t s = ctx->stream;
    // Dispatch on three streams so the classes overlap: B x B first on side
    // stream 0 (its tiles are the longest, and one launch holds at most one
    // workgroup per CU), the cell classes alternating between the main stream
    // and side stream 1. A class whose persistent workgroups cannot all be
    // resident waits for room; none waits for another class to finish.
    SH_HIP(hipEventRecord(ctx->fork_ev, s));
    for (hipStream_t a : ctx->aux) SH_HIP(hipStreamWaitEvent(a, ctx->fork_ev, 0));
    int rc = STORM_HIP_OK;
    hipStream_t sb = ctx->aux[0];
    if (timed) (void)hipEventRecord(ctx->ev[2], sb);
    const QItem* qbb = p->queue + (uint64_t)STORM_HIP_CELL_BB * p->qstride;
    // B x B items: whole tiles or K slices (at most qstride of them).
    const uint32_t gbb = (uint32_t)std::min<uint64_t>(
        p->qstride, (uint64_t)ctx->cu_count * bb_blocks_per_cu(p->tile, ctx->mfma));
    const bool thresholded = p->th.emits();
    // Per-item times (exploration and probe calls), per queue.
    auto ticks_of = [&](int q) -> unsigned long long* {
        return (with_ticks && p->ticks) ? p->ticks + (uint64_t)q * p->qstride : nullptr;
    };
    if (p->tile == 256 && thresholded)
        rc = launch_bb_mfma_queue(fp4, ctx->mfma, sb, p->d_corpora, qbb,
                                  &p->ctr->qcount[STORM_HIP_CELL_BB],
                                  &p->ctr->qhead[STORM_HIP_CELL_BB], p->thr_grid,
                                  ticks_of(STORM_HIP_CELL_BB), p->thr_buf);
    else if (p->tile == 256)
        rc = launch_bb_mfma_queue(fp4, ctx->mfma, sb, p->d_corpora, qbb,
                                  &p->ctr->qcount[STORM_HIP_CELL_BB],
                                  &p->ctr->qhead[STORM_HIP_CELL_BB], gbb,
                                  ticks_of(STORM_HIP_CELL_BB));
    else
        rc = launch_bb_valu_queue(sb, (int)p->tile, p->d_corpora, qbb,
                                  &p->ctr->qcount[STORM_HIP_CELL_BB],
                                  &p->ctr->qhead[STORM_HIP_CELL_BB], gbb,
                                  ticks_of(STORM_HIP_CELL_BB), thresholded);
    if (timed) (void)hipEventRecord(ctx->ev[3], sb);
    // Wide corpora's B x B tiles (int8) on side stream 1, ahead of the joins,
    // launched only when some corpus of the plan has them: an empty launch
    // still waits for room on the CUs and would hold the joins behind it.
    bool any_wide = false;
    for (storm_hip_corpus* c : p->cs) any_wide = any_wide || (c->n_bitmap_rows && bb_wide(c, o, p->tile, thresholded));
    if (rc == STORM_HIP_OK && p->tile == 256 && !thresholded && any_wide)
        rc = launch_bb_mfma_queue(false, ctx->mfma, ctx->aux[1], p->d_corpora,
                                  p->queue + (uint64_t)kQueueBBWide * p->qstride,
                                  &p->ctr->qcount[kQueueBBWide], &p->ctr->qhead[kQueueBBWide],
                                  gbb, ticks_of(kQueueBBWide));
    // Persistent grids: as many workgroups as fit on the chip at once. The
    // classes that usually carry the most work (B x S stripes, the R x R and
    // S x S joins) start at once on different streams: B x S first on the
    // main stream, the joins first on side stream 1; the per-pair classes
    // follow them.
    static const int kOrder[2][5] = {
        {STORM_HIP_CELL_BS, STORM_HIP_CELL_SR, STORM_HIP_CELL_SW, STORM_HIP_CELL_RW,
         STORM_HIP_CELL_CB},
        {STORM_HIP_CELL_RR, STORM_HIP_CELL_SS, STORM_HIP_CELL_BR, STORM_HIP_CELL_BW,
         STORM_HIP_CELL_WW}};
    static_assert(STORM_HIP_NCELL == 12, "every cell class but B x B and EMPTY in kOrder");
    for (int x = 0; x < 5 && rc == STORM_HIP_OK; ++x)
        for (int side = 0; side < 2 && rc == STORM_HIP_OK; ++side) {
            const int cell = kOrder[side][x];
            const uint32_t grid = (uint32_t)std::min<uint64_t>(
                p->qstride, (uint64_t)ctx->cu_count * cell_blocks_per_cu(cell));
            rc = launch_cell(cell, side ? ctx->aux[1] : s, grid, p->d_corpora, M,
                             p->queue + (uint64_t)cell * p->qstride, p->ctr, ticks_of(cell));
        }
    // Refined and per-pair tiles (REFINE and PER_PAIR calls only queue any).
    if (rc == STORM_HIP_OK &&
        (o.policy == STORM_HIP_POLICY_REFINE || o.policy == STORM_HIP_POLICY_PER_PAIR)) {
        const uint32_t grid = (uint32_t)std::min<uint64_t>(
            p->qstride, (uint64_t)ctx->cu_count * refine_blocks_per_cu());
        rc = launch_refine(s, grid, p->d_corpora, M,
                           p->queue + (uint64_t)kQueueRefine * p->qstride, p->ctr,
                           ticks_of(kQueueRefine));
    }
    // Tiles settled empty (zeros only) go last.
    if (rc == STORM_HIP_OK) {
        const uint32_t grid = (uint32_t)std::min<uint64_t>(
            p->qstride, (uint64_t)ctx->cu_count * cell_blocks_per_cu(STORM_HIP_CELL_EMPTY));
        rc = launch_cell(STORM_HIP_CELL_EMPTY, ctx->aux[1], grid, p->d_corpora, M,
                         p->queue + (uint64_t)STORM_HIP_CELL_EMPTY * p->qstride, p->ctr, nullptr);
    }
    // Join the side streams back, whatever happened (a capture must close).
    for (int a = 0; a < 2; ++a) {
        const hipError_t e1 = hipEventRecord(ctx->join_ev[a], ctx->aux[a]);
        const hipError_t e2 = hipStreamWaitEvent(s, ctx->join_ev[a], 0);
        if (rc == STORM_HIP_OK && (e1 != hipSuccess || e2 != hipSuccess)) {
            set_error("stream join failed: %s", hipGetErrorString(e1 != hipSuccess ? e1 : e2));
            rc = STORM_HIP_ERR_DEVICE;
        }
    }
    return rc;
}

// Workgroups each class shares a CU with while its queue drains: the
// divisor that turns an item's wall time into CU-time.
ProbeBpc probe_bpc(storm_hip_ctx* ctx, const Plan* p) {
    ProbeBpc b;
    const uint32_t cu = (uint32_t)std::max(1, ctx->cu_count);
    for (int c = 0; c < kNumCells; ++c) {
        const uint32_t per_cu = c == STORM_HIP_CELL_BB ? bb_blocks_per_cu(p->tile, ctx->mfma)
                                                       : cell_blocks_per_cu(c);
        const uint32_t grid = (uint32_t)std::min<uint64_t>(p->qstride, (uint64_t)cu * per_cu);
        b.v[c] = (float)std::max(1.0, std::min((double)per_cu, (double)grid / cu));
    }
    return b;
}

// The kernel sequence of one selecting call on the context's stream. PROBE
// runs two select-and-dispatch passes with the measurement between them.
int enqueue_selecting(storm_hip_ctx* ctx, Plan* p, const storm_hip_options& o, bool timed) {
    hipStream_t s = ctx->stream;
    const DevModel M = dev_model(ctx->model);
    const size_t n = p->cs.size();
    const bool probe = o.policy == STORM_HIP_POLICY_PROBE;
    SH_TRY(zero_words(s, p->ctr, sizeof(DevCounters)));
    SH_TRY(zero_words(s, p->sums, 8 * n));
    SH_TRY(zero_words(s, p->finish, 8 * n));
    if (p->th.n_hits) SH_TRY(zero_words(s, p->th.n_hits, 8));
    if (p->th.kind) SH_TRY(zero_words(s, p->th.buf, p->th.buf_bytes));
    if (probe) SH_TRY(zero_words(s, p->acc, sizeof(ProbeAcc)));
    hipLaunchKernelGGL(stamp_kernel, dim3(1), dim3(1), 0, s, p->finish + n);
    SH_HIP(hipGetLastError());
    if (timed) SH_HIP(hipEventRecord(ctx->ev[4], s));
    const uint32_t sel_policy = probe ? kPolicyProbe1 : o.policy;
    uint32_t max_words = 0;
    for (size_t k = 0; k < n; ++k) {
        storm_hip_corpus* c = p->cs[k];
        Work* w = p->works[k];
        SH_TRY(launch_tile_meta(s, p->h_corpora[k], w->tile, w->T, w->tm));
        const float* lc = nullptr;
        const uint8_t* lok = nullptr;
        if (o.policy == STORM_HIP_POLICY_LEARNED) {
            bool have = false;
            SH_TRY(learned_scores(ctx, w->tm, w->T, w->P, c->universe, w->learned_cost,
                                  w->learned_ok, &have));
            if (have) { lc = w->learned_cost; lok = w->learned_ok; }
        }
        // Corpora whose B x B tiles run as K slices add into a zeroed triangle.
        const uint32_t ks = p->h_corpora[k].bb_ksplit;
        if (ks > 1 && p->h_corpora[k].tri)
            SH_TRY(zero_words(s, p->h_corpora[k].tri, (uint64_t)c->n_rows * (c->n_rows - 1) / 2 * 4));
        SH_TRY(launch_select(s, w->tm, w->T, M, reps_of(c), sel_policy, o.fixed_cell, lc, lok,
                             w->decisions, p->queue, p->qstride, (uint32_t)k, p->tile,
                             p->th.thr, ks, bb_wide(c, o, p->tile, thresholded_plan(p)),
                             o.policy == kPolicyExplore ? p->explore_seed
                                                        : (probe ? kGolden : 0),
                             p->explore_share, p->d_sel, p->ctr, probe ? w->probe : nullptr,
                             p->acc, (uint32_t)(k + n), bs_wide_call(c, sel_policy, p->tile),
                             w->tzone, c->zone_ld));
        SH_TRY(launch_compose_rowmap(s, p->h_corpora[k].perm, c->d_bm_row, w->npad, w->rowmap));
        // Only corpora with bitmap rows can queue B x B tiles.
        if (c->n_bitmap_rows) max_words = std::max(max_words, c->n_words);
    }
    // fp4 with f32 accumulation is exact while one item covers fewer than
    // 2^18 words (counts < 2^24): 256-row B x B tiles run on the matrix cores,
    // fp4 unless int8 is asked for; wide universes split their tiles into K
    // slices that add into the output as integers, so they stay fp4 up to
    // kBBMaxSlices slices of 2^18 words; only tiles beyond that go to the int8
    // queue (in thresholded calls, which do not split K, every corpus takes
    // the one format the widest allows).
    const bool thr_call = thresholded_plan(p);
    const bool fp4 = o.bb_kernel != STORM_HIP_BB_MFMA && (!thr_call || max_words < (1u << 18));
    if (p->tile == 256)
        for (size_t k = 0; k < n; ++k) {
            storm_hip_corpus* c = p->cs[k];
            if (c->n_bitmap_rows)
                SH_TRY(bb_mfma_check(fp4 && !bb_wide(c, o, p->tile, thr_call),
                                     ctx->mfma.stage_words, c->n_words, c->ld,
                                     bb_item_words(c->n_words, p->h_corpora[k].bb_ksplit)));
        }
    if (timed) SH_HIP(hipEventRecord(ctx->ev[5], s));
    int rc = dispatch_queues(ctx, p, o, M, fp4, timed,
                             o.policy == kPolicyExplore || (probe && p->ticks));
    if (rc != STORM_HIP_OK || !probe) return rc;
    // Probe: fold pass 1's times into the classes, then commit the pending
    // tiles and run them.
    SH_TRY(launch_probe_reduce(s, p->d_corpora, (uint32_t)n, p->queue, p->qstride, p->ctr,
                               p->ticks, probe_bpc(ctx, p), p->acc));
    for (size_t k = 0; k < n; ++k) {
        storm_hip_corpus* c = p->cs[k];
        Work* w = p->works[k];
        SH_TRY(launch_select(s, w->tm, w->T, M, reps_of(c), kPolicyProbe2, o.fixed_cell, nullptr,
                             nullptr, w->decisions, p->queue, p->qstride, (uint32_t)k, p->tile,
                             p->th.thr, p->h_corpora[k].bb_ksplit,
                             bb_wide(c, o, p->tile, thresholded_plan(p)), 0, 0, p->d_sel, p->ctr,
                             w->probe, p->acc, (uint32_t)(k + n), false, w->tzone,
                             c->zone_ld));
    }
    return dispatch_queues(ctx, p, o, M, fp4, false, false);
}

// Run the plan (graph or eager), then read sums, counters, decisions, latencies.
int run_plan(storm_hip_ctx* ctx, Plan* p, const storm_hip_options& o, storm_hip_stats* st,
             uint64_t* sums, double* latency_ms) {
    hipStream_t s = ctx->stream;
    if (o.policy == STORM_HIP_POLICY_PROBE && !p->ticks)
        SH_HIP(hipMalloc(&p->ticks, sizeof(unsigned long long) * p->qstride * kNumQueues));
    if (std::memcmp(p->sel, p->sel_dev, sizeof p->sel) != 0) {
        SH_HIP(hipMemcpyAsync(p->d_sel, p->sel, sizeof p->sel, hipMemcpyHostToDevice, s));
        std::memcpy(p->sel_dev, p->sel, sizeof p->sel);
    }
    if (o.use_graph) {
        const uint32_t key[4] = {o.policy, o.fixed_cell, o.bb_kernel, ctx->model_version};
        if (!p->graph || std::memcmp(key, p->graph_key, sizeof key) != 0) {
            if (p->graph) { (void)hipGraphExecDestroy(p->graph); p->graph = nullptr; }
            hipGraph_t g;
            SH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue_selecting(ctx, p, o, false);
            hipError_t e = hipStreamEndCapture(s, &g);
            if (rc != STORM_HIP_OK) {
                if (e == hipSuccess) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) { set_error("graph capture: %s", hipGetErrorString(e)); return STORM_HIP_ERR_DEVICE; }
            e = hipGraphInstantiate(&p->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) { set_error("graph instantiate: %s", hipGetErrorString(e)); return STORM_HIP_ERR_DEVICE; }
            std::memcpy(p->graph_key, key, sizeof key);
        }
        SH_HIP(hipGraphLaunch(p->graph, s));
    } else {
        SH_TRY(enqueue_selecting(ctx, p, o, true));
    }
    const size_t n = p->cs.size();
    std::vector<unsigned long long> hs(n), hf(n + 1);
    DevCounters h;
    unsigned long long probe_st[2] = {0, 0};
    SH_HIP(hipMemcpyAsync(&h, p->ctr, sizeof h, hipMemcpyDeviceToHost, s));
    if (o.policy == STORM_HIP_POLICY_PROBE)
        SH_HIP(hipMemcpyAsync(probe_st, &p->acc->sampled, sizeof probe_st, hipMemcpyDeviceToHost, s));
    if (n) SH_HIP(hipMemcpyAsync(hs.data(), p->sums, 8 * n, hipMemcpyDeviceToHost, s));
    SH_HIP(hipMemcpyAsync(hf.data(), p->finish, 8 * (n + 1), hipMemcpyDeviceToHost, s));
    for (size_t k = 0; k < n; ++k) {
        storm_hip_corpus* c = p->cs[k];
        const Work* w = p->works[k];
        c->last_cells.resize(w->P);
        if (w->P) SH_HIP(hipMemcpyAsync(c->last_cells.data(), w->decisions, w->P, hipMemcpyDeviceToHost, s));
    }
    SH_HIP(hipStreamSynchronize(s));
    uint64_t total = 0, decisions = 0;
    for (size_t k = 0; k < n; ++k) {
        total += hs[k];
        if (sums) sums[k] = hs[k];
        // wall_clock64 runs at 100 MHz; a corpus with no tile keeps finish 0.
        if (latency_ms) latency_ms[k] = hf[k] > hf[n] ? (double)(hf[k] - hf[n]) * 1e-5 : 0.0;
        decisions += p->works[k]->P;
    }
    st->sum = total;
    for (int k = 0; k < kNumCells; ++k) {
        st->cell_pairs[k] = h.cell_pairs[k];
        st->cell_tiles[k] = h.cell_tiles[k];
    }
    st->skipped = h.skipped;
    st->fallback_tiles = h.fallback_tiles;
    st->decisions = decisions;
    st->bb_launches = (h.qcount[STORM_HIP_CELL_BB] ? 1 : 0) + (h.qcount[kQueueBBWide] ? 1 : 0);
    st->probe_tiles = probe_st[0];
    st->probe_switched = probe_st[1];
    if (!o.use_graph) {
        float ms = 0;
        hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]);
        st->ms_select = ms;
        hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]);
        st->ms_bb = ms;
    }
    return STORM_HIP_OK;
}

// One exploration call over the plan's corpora, then per non-empty tile pair
// its features, the cell it ran and that cell's cost: the summed CU-time of
// its work items (wall time over the class's workgroups per CU), in ns.
int explore_plan(storm_hip_ctx* ctx, Plan* p, uint32_t tile, uint64_t seed, double fraction,
                 uint32_t capacity, float* features, float* log2_ns, uint8_t* mask,
                 uint32_t* tile_index, uint32_t* corpus_index, uint32_t* n_out, uint64_t* sums) {
    hipStream_t s = ctx->stream;
    if (!p->ticks)
        SH_HIP(hipMalloc(&p->ticks, sizeof(unsigned long long) * p->qstride * kNumQueues));
    p->explore_seed = seed ? seed : kGolden;
    p->explore_share = (uint32_t)std::min(16777216.0, std::max(0.0, fraction) * 16777216.0);
    storm_hip_options o{};
    o.policy = kPolicyExplore;
    o.tile = tile;
    storm_hip_stats st{};
    const int rc = run_plan(ctx, p, o, &st, sums, nullptr);
    p->explore_seed = 0;
    SH_TRY(rc);
    DevCounters h;
    SH_HIP(hipMemcpy(&h, p->ctr, sizeof h, hipMemcpyDeviceToHost));
    const int F = STORM_HIP_MLP_FEATURES;
    const size_t n = p->cs.size();
    // Tile-pair index base of each corpus in one flat array of costs.
    std::vector<uint64_t> base(n + 1, 0);
    for (size_t k = 0; k < n; ++k) base[k + 1] = base[k] + p->works[k]->P;
    std::vector<double> cost(base[n], 0.0);
    std::vector<int8_t> cell_of(base[n], -1);
    const uint32_t cu = (uint32_t)std::max(1, ctx->cu_count);
    for (int q = 0; q < kNumQueues; ++q) {
        const uint32_t cnt = h.qcount[q];
        if (!cnt || q == STORM_HIP_CELL_EMPTY || q == kQueueRefine) continue;
        const int cell = q == kQueueBBWide ? STORM_HIP_CELL_BB : q;
        std::vector<QItem> items(cnt);
        std::vector<unsigned long long> ticks(cnt);
        SH_HIP(hipMemcpy(items.data(), p->queue + (uint64_t)q * p->qstride, cnt * sizeof(QItem),
                         hipMemcpyDeviceToHost));
        SH_HIP(hipMemcpy(ticks.data(), p->ticks + (uint64_t)q * p->qstride, cnt * 8,
                         hipMemcpyDeviceToHost));
        const uint32_t per_cu = cell == STORM_HIP_CELL_BB ? bb_blocks_per_cu(tile, ctx->mfma)
                                                          : cell_blocks_per_cu(cell);
        const uint32_t grid = (uint32_t)std::min<uint64_t>(p->qstride, (uint64_t)cu * per_cu);
        const double sharing = std::max(1.0, std::min((double)per_cu, (double)grid / cu));
        for (uint32_t e = 0; e < cnt; ++e) {
            const QItem& it = items[e];
            const uint32_t T = p->works[it.inst]->T;
            // A wide stripe serves tile pairs (bi, bj) and (bi, bj + 1): half
            // of its time to each.
            const bool wide = it.sub && (it.sub >> 24) == kWideSide;
            const double ns = (double)ticks[e] * 10.0 / sharing;   // 100 MHz clock -> ns
            for (uint32_t h = 0; h < (wide ? 2u : 1u); ++h) {
                const uint64_t t = base[it.inst] + tri_tile_id(it.bi, it.bj + h, T);
                cost[t] += wide ? ns / 2 : ns;
                cell_of[t] = (int8_t)cell;
            }
        }
    }
    uint32_t m = 0;
    for (size_t k = 0; k < n; ++k) {
        const Work* w = p->works[k];
        if (!w->P) continue;
        float* dX = nullptr;
        SH_HIP(hipMalloc(&dX, sizeof(float) * F * w->P));
        std::vector<float> X((size_t)F * w->P);
        int frc = launch_features(s, w->tm, w->T, w->P, p->cs[k]->universe, dX);
        if (frc == STORM_HIP_OK &&
            (hipMemcpyAsync(X.data(), dX, X.size() * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
             hipStreamSynchronize(s) != hipSuccess))
            frc = STORM_HIP_ERR_DEVICE;
        (void)hipFree(dX);
        SH_TRY(frc);
        for (uint64_t t = 0; t < w->P && m < capacity; ++t) {
            const uint64_t g = base[k] + t;
            if (cell_of[g] < 0) continue;
            std::memcpy(features + (size_t)m * F, &X[(size_t)t * F], F * 4);
            for (int c = 0; c < 10; ++c) {
                mask[(size_t)m * 10 + c] = c == cell_of[g];
                log2_ns[(size_t)m * 10 + c] =
                    c == cell_of[g] ? (float)std::log2(std::max(cost[g], 10.0)) : 0.f;
            }
            tile_index[m] = (uint32_t)t;
            if (corpus_index) corpus_index[m] = (uint32_t)k;
            ++m;
        }
    }
    *n_out = m;
    return STORM_HIP_OK;
}

int check_tile(uint32_t tile) {
    if (tile != 64 && tile != 128 && tile != 256) {
        set_error("tile must be 64, 128 or 256 rows (got %u)", tile);
        return STORM_HIP_ERR_INVALID;
    }
    return STORM_HIP_OK;
}

}  // namespace

void general_free(storm_hip_corpus* c) {
    free_plan(static_cast<Plan*>(c->plan));
    c->plan = nullptr;
    free_work(static_cast<Work*>(c->work));
    c->work = nullptr;
    if (c->d_hits) (void)hipFree(c->d_hits);
    if (c->d_n_hits) (void)hipFree(c->d_n_hits);
    if (c->d_emit) (void)hipFree(c->d_emit);
    c->d_emit = nullptr;
    c->emit_bytes = 0;
    c->d_hits = nullptr;
    c->d_n_hits = nullptr;
    if (!c->dense) general_free_arenas(c);
}

int general_allpairs(storm_hip_ctx* ctx, storm_hip_corpus* c, const storm_hip_options* opt,
                     storm_hip_stats* st, unsigned long long* d_sum) {
    const storm_hip_options& o = *opt;
    const uint32_t tile = o.tile ? o.tile : 64;
    SH_TRY(check_tile(tile));
    hipStream_t s = ctx->stream;
    const bool by_position = o.out_mode == STORM_HIP_OUT_UPPER_TRI_DENSITY;
    uint32_t* tri = (o.out_mode != STORM_HIP_OUT_SUM) ? c->d_tri : nullptr;
    if (c->n_rows < 2) return STORM_HIP_OK;

    if (o.policy == STORM_HIP_POLICY_ALL_BITMAP) {
        // Every pair B x B, no span test (storm_allpairs.cpp: AllBitmap),
        // bitmaps materialised on the device where missing. The walk takes
        // the density order, in which the bitmap store is laid out (a tile's
        // rows are one contiguous run of the store); the counts are the same
        // in any order, written for rows perm[i], perm[j] (original order)
        // or at the positions themselves (position-packed triangle).
        SH_TRY(general_materialize(ctx, c));
        Work* w = nullptr;
        SH_TRY(ensure_bitmap_rowmap(c, &w));
        SH_TRY(launch_compose_rowmap(s, c->d_perm, c->d_bm_row, w->npad, w->rowmap));
        SH_HIP(hipEventRecord(ctx->ev[2], s));
        SH_TRY(launch_bb_dense(o.bb_kernel, ctx->mfma, s, c->d_bitmap, c->ld, c->n_words, w->rowmap,
                               c->n_rows, tri, d_sum, (uint32_t)ctx->cu_count,
                               by_position ? nullptr : c->d_perm));
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
    Plan* p = static_cast<Plan*>(c->plan);
    Thresh out;
    out.tri_density = by_position;
    const int rc = ensure_plan(p, {c}, {tri}, tile, out);
    c->plan = p;
    SH_TRY(rc);
    uint64_t sum = 0;
    SH_TRY(run_plan(ctx, p, o, st, &sum, nullptr));
    // Fold into the caller's accumulator.
    SH_HIP(hipMemcpyAsync(d_sum, p->sums, 8, hipMemcpyDeviceToDevice, s));
    return STORM_HIP_OK;
}

namespace {

// 256-row B x B tiles of calls that emit per pair run on one persistent
// matrix-core block per CU, each with its own 256 x 256 count buffer.
int ensure_thr_buf(storm_hip_ctx* ctx, Plan* p) {
    if (p->tile != 256 || p->thr_buf) return STORM_HIP_OK;
    p->thr_grid = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(p->totalP, 1),
                                               (uint64_t)ctx->cu_count);
    SH_HIP(hipMalloc(&p->thr_buf, (size_t)p->thr_grid * 256 * 256 * 4));
    return STORM_HIP_OK;
}

// One thresholded call into the corpus's pair buffer; returns the number of
// reportable pairs (which may exceed the buffer's capacity).
int threshold_pass(storm_hip_ctx* ctx, storm_hip_corpus* c, const storm_hip_options& o,
                   uint32_t tile, double t, storm_hip_stats* st, unsigned long long* n_hits) {
    Thresh th;
    th.thr = t;
    th.hits = static_cast<PairHit*>(c->d_hits);
    th.n_hits = c->d_n_hits;
    th.cap = c->hit_cap;
    Plan* p = static_cast<Plan*>(c->plan);
    const int rc = ensure_plan(p, {c}, {nullptr}, tile, th);
    c->plan = p;
    SH_TRY(rc);
    SH_TRY(ensure_thr_buf(ctx, p));
    uint64_t sum = 0;
    SH_TRY(run_plan(ctx, p, o, st, &sum, nullptr));
    SH_HIP(hipMemcpyAsync(n_hits, c->d_n_hits, 8, hipMemcpyDeviceToHost, ctx->stream));
    SH_HIP(hipStreamSynchronize(ctx->stream));
    return STORM_HIP_OK;
}

int ensure_emit(storm_hip_corpus* c, uint64_t bytes) {
    if (c->d_emit && c->emit_bytes == bytes) return STORM_HIP_OK;
    if (c->d_emit) { SH_HIP(hipFree(c->d_emit)); c->d_emit = nullptr; c->emit_bytes = 0; }
    SH_HIP(hipMalloc((void**)&c->d_emit, std::max<uint64_t>(bytes, 4)));
    c->emit_bytes = bytes;
    return STORM_HIP_OK;
}

// Options of a band or cross call: a general corpus and a selecting policy;
// the all-bitmap policy runs every tile B x B over materialised bitmaps.
int range_call_options(storm_hip_ctx* ctx, storm_hip_corpus* c, const storm_hip_options* opt,
                       storm_hip_options* o, uint32_t* tile) {
    *o = storm_hip_options{};
    if (opt) *o = *opt;
    if (o->policy > STORM_HIP_POLICY_PER_PAIR || o->bb_kernel > STORM_HIP_BB_MFMA_FP4 ||
        o->fixed_cell >= STORM_HIP_NCELL)
        return STORM_HIP_ERR_INVALID;
    if (c->dense) {
        set_error("band and cross calls take a general corpus");
        return STORM_HIP_ERR_UNSUPPORTED;
    }
    if (o->policy == STORM_HIP_POLICY_EXHAUSTIVE) {
        set_error("band and cross calls take a selecting policy or all-bitmap");
        return STORM_HIP_ERR_UNSUPPORTED;
    }
    *tile = o->tile ? o->tile : 64;
    SH_TRY(check_tile(*tile));
    o->tile = *tile;
    if (o->policy == STORM_HIP_POLICY_ALL_BITMAP) {
        SH_TRY(general_materialize(ctx, c));
        o->policy = STORM_HIP_POLICY_FIXED;
        o->fixed_cell = STORM_HIP_CELL_BB;
    }
    return STORM_HIP_OK;
}

// Run the corpus's plan for `th` over tile pairs sel, timing the call.
int run_range(storm_hip_ctx* ctx, storm_hip_corpus* c, const storm_hip_options& o, uint32_t tile,
              const Thresh& th, const uint32_t sel[4], storm_hip_stats* st) {
    Plan* p = static_cast<Plan*>(c->plan);
    const int rc = ensure_plan(p, {c}, {nullptr}, tile, th);
    c->plan = p;
    SH_TRY(rc);
    if (th.emits()) SH_TRY(ensure_thr_buf(ctx, p));
    std::memcpy(p->sel, sel, sizeof p->sel);
    SH_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    uint64_t sum = 0;
    const int rr = run_plan(ctx, p, o, st, &sum, nullptr);
    static const uint32_t kAll[4] = {0, 0xFFFFFFFFu, 0, 0xFFFFFFFFu};
    std::memcpy(p->sel, kAll, sizeof p->sel);
    SH_TRY(rr);
    SH_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    SH_HIP(hipEventSynchronize(ctx->ev[1]));
    float ms = 0;
    hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    st->ms_device = ms;
    return STORM_HIP_OK;
}

int grow_hits(storm_hip_corpus* c, uint64_t cap) {
    if (c->d_hits) { SH_HIP(hipFree(c->d_hits)); c->d_hits = nullptr; c->hit_cap = 0; }
    SH_HIP(hipMalloc(&c->d_hits, cap * sizeof(PairHit)));
    c->hit_cap = cap;
    return STORM_HIP_OK;
}

}  // namespace

}  // namespace stormhip

using namespace stormhip;

struct storm_hip_batch {
    storm_hip_ctx* ctx = nullptr;
    std::vector<storm_hip_corpus*> cs;
    Plan* plan = nullptr;
};

extern "C" {

int storm_hip_batch_create(storm_hip_ctx* ctx, storm_hip_corpus* const* corpora, uint32_t n,
                           storm_hip_batch** out) {
    if (!ctx || !out || (!corpora && n)) return STORM_HIP_ERR_INVALID;
    for (uint32_t k = 0; k < n; ++k)
        if (!corpora[k] || corpora[k]->ctx != ctx || corpora[k]->dense) {
            set_error("batch member %u is not a general corpus of this context", k);
            return STORM_HIP_ERR_INVALID;
        }
    storm_hip_batch* b = new (std::nothrow) storm_hip_batch();
    if (!b) return STORM_HIP_ERR_NOMEM;
    b->ctx = ctx;
    b->cs.assign(corpora, corpora + n);
    *out = b;
    return STORM_HIP_OK;
}

void storm_hip_batch_free(storm_hip_batch* b) {
    if (!b) return;
    if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
    free_plan(b->plan);
    delete b;
}

int storm_hip_batch_allpairs(storm_hip_ctx* ctx, storm_hip_batch* b, const storm_hip_options* opt,
                             storm_hip_stats* st, uint64_t* sums, double* latency_ms) {
    if (!ctx || !b || b->ctx != ctx || !opt) return STORM_HIP_ERR_INVALID;
    const storm_hip_options& o = *opt;
    if (o.policy > STORM_HIP_POLICY_PER_PAIR || o.bb_kernel > STORM_HIP_BB_MFMA_FP4 ||
        o.fixed_cell >= STORM_HIP_NCELL || o.out_mode > STORM_HIP_OUT_UPPER_TRI_DENSITY)
        return STORM_HIP_ERR_INVALID;
    if (o.policy == STORM_HIP_POLICY_ALL_BITMAP || o.policy == STORM_HIP_POLICY_EXHAUSTIVE) {
        set_error("batched calls take a selecting policy (not all-bitmap or exhaustive)");
        return STORM_HIP_ERR_UNSUPPORTED;
    }
    const uint32_t tile = o.tile ? o.tile : 64;
    SH_TRY(check_tile(tile));
    SH_HIP(hipSetDevice(ctx->device));
    storm_hip_stats local{};
    storm_hip_stats& s = st ? *st : local;
    std::memset(&s, 0, sizeof s);
    std::vector<uint32_t*> tris(b->cs.size(), nullptr);
    for (size_t k = 0; k < b->cs.size(); ++k) {
        storm_hip_corpus* c = b->cs[k];
        const uint64_t n = c->n_rows, pairs = n ? n * (n - 1) / 2 : 0;
        s.pairs += pairs;
        c->tri_valid = false;
        if (o.out_mode != STORM_HIP_OUT_SUM) {
            if (c->tri_count != pairs) {
                if (c->d_tri) { SH_HIP(hipFree(c->d_tri)); c->d_tri = nullptr; }
                if (pairs) SH_HIP(hipMalloc(&c->d_tri, pairs * 4));
                c->tri_count = pairs;
            }
            tris[k] = c->d_tri;
        }
    }
    SH_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    Thresh out;
    out.tri_density = o.out_mode == STORM_HIP_OUT_UPPER_TRI_DENSITY;
    SH_TRY(ensure_plan(b->plan, b->cs, tris, tile, out));
    SH_TRY(run_plan(ctx, b->plan, o, &s, sums, latency_ms));
    SH_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    SH_HIP(hipEventSynchronize(ctx->ev[1]));
    float ms = 0;
    hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    s.ms_device = ms;
    for (size_t k = 0; k < b->cs.size(); ++k) {
        b->cs[k]->tri_valid = tris[k] != nullptr;
        b->cs[k]->tri_density = out.tri_density && tris[k] && !b->cs[k]->dense;
    }
    return STORM_HIP_OK;
}

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
    SH_TRY(check_tile(tile));
    *n_sampled = 0;
    if (c->n_rows < 2 || max_tiles == 0) return STORM_HIP_OK;
    SH_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    // One analytic pass finds the proven-empty tiles and fills the tile metas.
    Plan* p = static_cast<Plan*>(c->plan);
    int rc = ensure_plan(p, {c}, {nullptr}, tile);
    c->plan = p;
    SH_TRY(rc);
    storm_hip_options o{};
    o.policy = STORM_HIP_POLICY_ANALYTIC;
    o.tile = tile;
    storm_hip_stats st{};
    SH_TRY(run_plan(ctx, p, o, &st, nullptr, nullptr));
    const Work* w = p->works[0];
    const int F = STORM_HIP_MLP_FEATURES;
    float* dX = nullptr;
    SH_HIP(hipMalloc(&dX, sizeof(float) * F * w->P));
    std::vector<float> X((size_t)F * w->P);
    std::vector<TileMeta> tm(w->T);
    rc = launch_features(s, w->tm, w->T, w->P, c->universe, dX);
    if (rc == STORM_HIP_OK) {
        (void)hipMemcpyAsync(X.data(), dX, X.size() * 4, hipMemcpyDeviceToHost, s);
        (void)hipMemcpyAsync(tm.data(), w->tm, tm.size() * sizeof(TileMeta), hipMemcpyDeviceToHost, s);
        if (hipStreamSynchronize(s) != hipSuccess) rc = STORM_HIP_ERR_DEVICE;
    }
    (void)hipFree(dX);
    if (rc != STORM_HIP_OK) return rc;
    const std::vector<uint8_t>& dec = c->last_cells;
    // Seeded sample of the non-empty tile pairs.
    std::vector<uint32_t> cand;
    for (uint64_t t = 0; t < w->P; ++t)
        if (dec[t] != STORM_HIP_CELL_EMPTY) cand.push_back((uint32_t)t);
    uint64_t rs = seed ? seed : kGolden;
    for (size_t i = cand.size(); i > 1; --i) {
        rs += kGolden;
        std::swap(cand[i - 1], cand[splitmix_mix(rs) % i]);
    }
    // The rest of the shuffled order: tiles whose items fill the workgroup
    // slots while the sampled ones are timed.
    std::vector<uint32_t> rest;
    if (cand.size() > max_tiles) {
        rest.assign(cand.begin() + max_tiles, cand.end());
        cand.resize(max_tiles);
    }
    std::sort(cand.begin(), cand.end());
    const uint32_t S = (uint32_t)cand.size();
    std::vector<QItem> items(S);
    std::vector<uint32_t> elig(S);
    std::vector<double> scale(S, 1.0);
    const uint32_t reps = reps_of(c);
    const DevModel M = dev_model(ctx->model);
    // The per-pair cells are timed on a kSampleRows x kSampleRows block of
    // median-density rows of each tile and the time is scaled by the tile's
    // pair count over the block's (their cost is per pair).
    constexpr uint32_t kSampleRows = 64;
    for (uint32_t k = 0; k < S; ++k) {
        uint32_t bi, bj;
        tri_tile(cand[k], w->T, bi, bj);
        // The cells the learned selector may rank on this tile.
        elig[k] = learned_candidates(M, tm[bi], tm[bj], eligible_mask(tm[bi], tm[bj], reps));
        tile_index[k] = cand[k];
        std::memcpy(features + (size_t)k * F, &X[(size_t)cand[k] * F], F * 4);
        items[k] = QItem{0, bi, bj, 0};
        if (tile > kSampleRows) {
            // The middle block of each side (the same block on a diagonal tile).
            const uint32_t ri = tm[bi].rows, rj = tm[bj].rows, mid = tile / kSampleRows / 2;
            uint32_t xi = std::min(mid, ri ? (ri - 1) / kSampleRows : 0);
            uint32_t xj = std::min(mid, rj ? (rj - 1) / kSampleRows : 0);
            if (bi == bj) xi = xj = std::min(xi, xj);
            items[k].sub = make_sub(kSampleRows, xi, xj);
            const double si = std::min<double>(kSampleRows, ri - xi * kSampleRows);
            const double sj = std::min<double>(kSampleRows, rj - xj * kSampleRows);
            const double full = bi == bj ? (double)ri * (ri - 1) / 2 : (double)ri * rj;
            const double part = bi == bj ? si * (si - 1) / 2 : si * sj;
            scale[k] = part > 0 ? full / part : 1.0;
        }
    }
    // Time every eligible cell on its sampled tiles, each class alone, as a
    // persistent launch like the production one. A tile's cost is its CU time:
    // its wall-clock latency divided by the number of workgroups sharing its
    // CU during the measurement (cells that fit several workgroups per CU
    // overlap them; the matrix-core B x B holds a CU alone).
    unsigned long long* dticks = nullptr;
    SH_HIP(hipMalloc(&dticks, sizeof(unsigned long long) * std::max<uint32_t>(S, 1)));
    std::vector<unsigned long long> ticks(S);
    std::vector<double> best((size_t)S * 10, 0.0);
    std::vector<uint8_t> has((size_t)S * 10, 0);
    // B x B items as production queues them: whole 256-row tiles, or their K
    // slices on wide universes (fp4 while each slice stays exact in f32).
    const uint32_t bb_ks = tile == 256 ? std::max(1u, p->h_corpora[0].bb_ksplit) : 1u;
    const bool fp4 = bb_item_words(c->n_words, bb_ks) < (1u << 18);
    if (tile == 256 && c->n_bitmap_rows)
        SH_TRY(bb_mfma_check(fp4, ctx->mfma.stage_words, c->n_words, c->ld,
                             bb_item_words(c->n_words, bb_ks)));
    const uint32_t cu = (uint32_t)std::max(1, ctx->cu_count);
    // The block-join cells (S x S, R x R) and the B x S stripes are timed on
    // every block or stripe of the tile, queued as the selector queues them:
    // their cost per block follows the rows' sizes, which vary across a tile.
    // The per-pair cells take one block and scale by the pair count.
    auto whole_tile = [&](int cell) {
        return tile > kSplitRows && (cell == STORM_HIP_CELL_SS || cell == STORM_HIP_CELL_RR ||
                                     cell == STORM_HIP_CELL_BS);
    };
    constexpr uint32_t kFiller = 0xFFFFFFFFu;   // an item that loads the GPU, not timed
    size_t next_rest = 0;
    for (int cell = 0; cell < 10 && rc == STORM_HIP_OK; ++cell) {
        std::vector<QItem> q;
        std::vector<uint32_t> slot_of;
        std::vector<double> factor;
        // The items of tile (bi, bj) as this cell runs them, for sampled tile
        // k (or a filler).
        auto add_items = [&](uint32_t bi, uint32_t bj, uint32_t k, const QItem& single,
                             double single_factor) {
            if (whole_tile(cell)) {
                const TileMeta& a = tm[bi];
                const TileMeta& b = tm[bj];
                const uint32_t ni = (a.rows + kSplitRows - 1) / kSplitRows;
                const uint32_t nj = (b.rows + kSplitRows - 1) / kSplitRows;
                if (cell == STORM_HIP_CELL_BS) {
                    const bool pa = bs_probe_a(a, b);
                    for (uint32_t x = 0; x < (pa ? ni : nj); ++x) {
                        q.push_back(QItem{0, bi, bj,
                                          pa ? make_sub(kSplitRows, x, kWholeSide)
                                             : make_sub(kSplitRows, kWholeSide, x)});
                        slot_of.push_back(k);
                        factor.push_back(1.0);
                    }
                } else {
                    for (uint32_t x = 0; x < ni; ++x)
                        for (uint32_t y = bi == bj ? x : 0; y < nj; ++y) {
                            q.push_back(QItem{0, bi, bj, make_sub(kSplitRows, x, y)});
                            slot_of.push_back(k);
                            factor.push_back(1.0);
                        }
                }
            } else if (cell == STORM_HIP_CELL_BB) {
                // Whole tiles, or all K slices of the tile (their times add).
                for (uint32_t x = 0; x < bb_ks; ++x) {
                    q.push_back(QItem{single.inst, bi, bj, bb_ks > 1 ? (x | (bb_ks << 16)) : 0u});
                    slot_of.push_back(k);
                    factor.push_back(1.0);
                }
            } else {
                q.push_back(single);
                slot_of.push_back(k);
                factor.push_back(single_factor);
            }
        };
        for (uint32_t k = 0; k < S; ++k)
            if (elig[k] >> cell & 1) add_items(items[k].bi, items[k].bj, k, items[k], scale[k]);
        if (q.empty()) continue;
        const uint32_t per_cu = cell == STORM_HIP_CELL_BB ? bb_blocks_per_cu(tile, ctx->mfma)
                                                          : cell_blocks_per_cu(cell);
        // Production launches keep every CU at full occupancy, and a
        // latency-bound cell's time per item under that load is not its time
        // alone on a CU. Free workgroup slots are filled with the items of
        // unsampled tiles (other rows, so the caches are as cold as in
        // production; not timed); where there are too few, the sampled items
        // are repeated and each costs the mean over its copies.
        const size_t base = q.size();
        const size_t slots = (size_t)cu * per_cu;
        const size_t cap = std::min<size_t>(p->qstride, std::max(slots, base));
        // Fillers are drawn from the rest of the shuffled tile order: any rows
        // of the corpus, so the mix of sizes and the cache misses are those of
        // a production launch.
        for (size_t tried = 0; q.size() < slots && tried < rest.size(); ++tried) {
            const uint32_t t = rest[next_rest++ % rest.size()];
            uint32_t bi, bj;
            tri_tile(t, w->T, bi, bj);
            if (!(eligible_mask(tm[bi], tm[bj], reps) >> cell & 1)) continue;
            const size_t before = q.size();
            add_items(bi, bj, kFiller, QItem{0, bi, bj, items[0].sub}, 0.0);
            if (q.size() > cap) {   // keep within the queue
                q.resize(before);
                slot_of.resize(before);
                factor.resize(before);
                break;
            }
        }
        size_t copies = q.size() >= slots ? 1 : std::max<size_t>(1, (slots + base - 1) / base);
        copies = std::min<size_t>(copies, std::max<size_t>(1, p->qstride / std::max<size_t>(q.size(), 1)));
        for (size_t c = 1; c < copies; ++c)
            for (size_t e = 0; e < base; ++e) {
                q.push_back(q[e]);
                slot_of.push_back(slot_of[e]);
                factor.push_back(factor[e] / (double)copies);
            }
        for (size_t e = 0; e < base; ++e) factor[e] /= (double)copies;
        if (q.size() > ticks.size()) {
            (void)hipFree(dticks);
            dticks = nullptr;
            ticks.resize(q.size());
            if (hipMalloc(&dticks, sizeof(unsigned long long) * q.size()) != hipSuccess) {
                rc = STORM_HIP_ERR_NOMEM;
                break;
            }
        }
        if (q.size() > p->qstride) { rc = STORM_HIP_ERR_INVALID; break; }
        QItem* dq = p->queue + (uint64_t)cell * p->qstride;
        const uint32_t grid = (uint32_t)std::min<uint64_t>(q.size(), (uint64_t)cu * per_cu);
        const double sharing = std::max(1.0, std::min((double)per_cu, (double)grid / cu));
        (void)hipMemcpyAsync(dq, q.data(), q.size() * sizeof(QItem), hipMemcpyHostToDevice, s);
        std::vector<double> sum(S);
        for (uint32_t r = 0; r < std::max(1u, repeats) && rc == STORM_HIP_OK; ++r) {
            DevCounters h{};
            h.qcount[cell] = (uint32_t)q.size();
            (void)hipMemcpyAsync(p->ctr, &h, sizeof h, hipMemcpyHostToDevice, s);
            if (cell == STORM_HIP_CELL_BB) {
                if (tile == 256)
                    rc = launch_bb_mfma_queue(fp4, ctx->mfma, s, p->d_corpora, dq,
                                              &p->ctr->qcount[cell], &p->ctr->qhead[cell], grid,
                                              dticks);
                else
                    rc = launch_bb_valu_queue(s, (int)tile, p->d_corpora, dq, &p->ctr->qcount[cell],
                                              &p->ctr->qhead[cell], grid, dticks);
            } else {
                rc = launch_cell(cell, s, grid, p->d_corpora, M, dq, p->ctr, dticks);
            }
            if (rc != STORM_HIP_OK) break;
            (void)hipMemcpyAsync(ticks.data(), dticks, q.size() * 8, hipMemcpyDeviceToHost, s);
            if (hipStreamSynchronize(s) != hipSuccess) { rc = STORM_HIP_ERR_DEVICE; break; }
            std::fill(sum.begin(), sum.end(), 0.0);
            // 100 MHz wall clock -> ns of CU time, summed over the tile's items.
            for (size_t e = 0; e < q.size(); ++e)
                if (slot_of[e] != kFiller)
                    sum[slot_of[e]] += (double)ticks[e] * 10.0 / sharing * factor[e];
            for (size_t e = 0; e < q.size(); ++e) {
                const uint32_t k = slot_of[e];
                if (k == kFiller) continue;
                if (!has[(size_t)k * 10 + cell] || sum[k] < best[(size_t)k * 10 + cell])
                    best[(size_t)k * 10 + cell] = sum[k];
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
    return storm_hip_calibrate_samples(ctx, c, tile, n, Y.data(), Mk.data(), idx.data(), out);
}

int storm_hip_selector_explore(storm_hip_ctx* ctx, storm_hip_corpus* c, uint32_t tile,
                               uint64_t seed, double fraction, uint32_t capacity, float* features,
                               float* log2_ns, uint8_t* mask, uint32_t* tile_index,
                               uint32_t* n_sampled, uint64_t* sum) {
    if (!ctx || !c || c->ctx != ctx || c->dense || !n_sampled) return STORM_HIP_ERR_INVALID;
    if (capacity && (!features || !log2_ns || !mask || !tile_index)) return STORM_HIP_ERR_NULL_VALUES;
    SH_TRY(check_tile(tile));
    *n_sampled = 0;
    if (c->n_rows < 2) return STORM_HIP_OK;
    SH_HIP(hipSetDevice(ctx->device));
    Plan* p = static_cast<Plan*>(c->plan);
    const int rc = ensure_plan(p, {c}, {nullptr}, tile);
    c->plan = p;
    SH_TRY(rc);
    return explore_plan(ctx, p, tile, seed, fraction, capacity, features, log2_ns, mask,
                        tile_index, nullptr, n_sampled, sum);
}

int storm_hip_batch_explore(storm_hip_ctx* ctx, storm_hip_batch* b, uint32_t tile, uint64_t seed,
                            double fraction, uint32_t capacity, float* features, float* log2_ns,
                            uint8_t* mask,
                            uint32_t* tile_index, uint32_t* corpus_index, uint32_t* n_sampled,
                            uint64_t* sums) {
    if (!ctx || !b || b->ctx != ctx || !n_sampled) return STORM_HIP_ERR_INVALID;
    if (capacity && (!features || !log2_ns || !mask || !tile_index || !corpus_index))
        return STORM_HIP_ERR_NULL_VALUES;
    SH_TRY(check_tile(tile));
    *n_sampled = 0;
    SH_HIP(hipSetDevice(ctx->device));
    std::vector<uint32_t*> tris(b->cs.size(), nullptr);
    SH_TRY(ensure_plan(b->plan, b->cs, tris, tile));
    return explore_plan(ctx, b->plan, tile, seed, fraction, capacity, features, log2_ns, mask,
                        tile_index, corpus_index, n_sampled, sums);
}

int storm_hip_calibrate_samples(storm_hip_ctx* ctx, storm_hip_corpus* c, uint32_t tile,
                                uint32_t n, const float* log2_ns, const uint8_t* mask,
                                const uint32_t* tile_index, storm_hip_cost_model* out) {
    if (!c) return STORM_HIP_ERR_INVALID;
    std::vector<uint32_t> which(n, 0);
    return storm_hip_calibrate_corpora(ctx, &c, 1, tile, n, log2_ns, mask, tile_index,
                                       which.data(), out);
}

int storm_hip_calibrate_corpora(storm_hip_ctx* ctx, storm_hip_corpus* const* corpora,
                                uint32_t n_corpora, uint32_t tile, uint32_t n,
                                const float* log2_ns, const uint8_t* mask,
                                const uint32_t* tile_index, const uint32_t* corpus_index,
                                storm_hip_cost_model* out) {
    if (!ctx || !corpora || n_corpora == 0) return STORM_HIP_ERR_INVALID;
    for (uint32_t q = 0; q < n_corpora; ++q)
        if (!corpora[q] || corpora[q]->ctx != ctx || corpora[q]->dense) return STORM_HIP_ERR_INVALID;
    if (n && (!log2_ns || !mask || !tile_index || !corpus_index)) return STORM_HIP_ERR_NULL_VALUES;
    SH_TRY(check_tile(tile));
    if (n == 0) {
        set_error("calibrate: no timed tile pair");
        return STORM_HIP_ERR_INVALID;
    }
    SH_HIP(hipSetDevice(ctx->device));
    // Each corpus's tile metadata at this tile width.
    std::vector<std::vector<TileMeta>> tms(n_corpora);
    std::vector<uint64_t> Ps(n_corpora, 0);
    std::vector<uint32_t> Ts(n_corpora, 0);
    std::vector<bool> used(n_corpora, false);
    for (uint32_t k = 0; k < n; ++k) {
        if (corpus_index[k] >= n_corpora) {
            set_error("calibrate: corpus index %u out of range (%u corpora)", corpus_index[k],
                      n_corpora);
            return STORM_HIP_ERR_RANGE;
        }
        used[corpus_index[k]] = true;
    }
    for (uint32_t q = 0; q < n_corpora; ++q) {
        if (!used[q]) continue;
        storm_hip_corpus* c = corpora[q];
        Plan* plan = static_cast<Plan*>(c->plan);
        const int prc = ensure_plan(plan, {c}, {nullptr}, tile);
        c->plan = plan;
        SH_TRY(prc);
        const Work* w = plan->works[0];
        SH_TRY(launch_tile_meta(ctx->stream, plan->h_corpora[0], w->tile, w->T, w->tm));
        tms[q].resize(w->T);
        SH_HIP(hipMemcpyAsync(tms[q].data(), w->tm, tms[q].size() * sizeof(TileMeta),
                              hipMemcpyDeviceToHost, ctx->stream));
        Ps[q] = w->P;
        Ts[q] = w->T;
    }
    SH_HIP(hipStreamSynchronize(ctx->stream));
    const float* Y = log2_ns;
    const uint8_t* Mk = mask;
    const uint32_t* idx = tile_index;
    for (uint32_t k = 0; k < n; ++k)
        if (idx[k] >= Ps[corpus_index[k]]) {
            set_error("calibrate: tile index %u out of range (%llu tile pairs)", idx[k],
                      (unsigned long long)Ps[corpus_index[k]]);
            return STORM_HIP_ERR_RANGE;
        }
    DeviceCostModel m = ctx->model;
    bool sampled[10] = {};
    for (int p = 0; p < 10; ++p) {
        std::vector<double> r;
        for (uint32_t k = 0; k < n; ++k) {
            if (!Mk[(size_t)k * 10 + p]) continue;
            const std::vector<TileMeta>& tm = tms[corpus_index[k]];
            uint32_t bi, bj;
            tri_tile(idx[k], Ts[corpus_index[k]], bi, bj);
            const double na = tm[bi].rows, nb = tm[bj].rows;
            const double pairs = bi == bj ? na * (na - 1) / 2 : na * nb;
            const double units = std::max(1.0, host_work_units(p, tm[bi], tm[bj], m.br_margin));
            if (pairs < 1) continue;
            r.push_back(std::exp2((double)Y[(size_t)k * 10 + p]) / (pairs * units));
        }
        if (r.empty()) continue;
        std::nth_element(r.begin(), r.begin() + r.size() / 2, r.end());
        m.ns_per_unit[p] = r[r.size() / 2];
        sampled[p] = true;
    }
    // Cells this corpus never ran keep their previous rate, rescaled by the
    // median ratio new/old over the cells it did run (so every rate is on the
    // device's scale and none is dropped from the candidate set).
    std::vector<double> ratio;
    for (int p = 0; p < 10; ++p)
        if (sampled[p] && ctx->model.ns_per_unit[p] > 0)
            ratio.push_back(m.ns_per_unit[p] / ctx->model.ns_per_unit[p]);
    if (ratio.empty()) {
        set_error("calibrate: no cell could be timed on this corpus");
        return STORM_HIP_ERR_INVALID;
    }
    std::nth_element(ratio.begin(), ratio.begin() + ratio.size() / 2, ratio.end());
    const double scale = ratio[ratio.size() / 2];
    for (int p = 0; p < 10; ++p)
        if (!sampled[p]) m.ns_per_unit[p] = ctx->model.ns_per_unit[p] * scale;
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

int storm_hip_allpairs_threshold(storm_hip_ctx* ctx, storm_hip_corpus* c,
                                 const storm_hip_options* opt, double t, storm_hip_stats* st,
                                 uint64_t* n_pairs) {
    if (!ctx || !c || c->ctx != ctx || !n_pairs) return STORM_HIP_ERR_INVALID;
    *n_pairs = 0;
    if (!(t > 0.0 && t <= 1.0)) {
        set_error("Jaccard threshold must be in (0, 1] (got %g)", t);
        return STORM_HIP_ERR_INVALID;
    }
    storm_hip_options o{};
    if (opt) o = *opt;
    if (o.policy > STORM_HIP_POLICY_PER_PAIR || o.bb_kernel > STORM_HIP_BB_MFMA_FP4 ||
        o.fixed_cell >= STORM_HIP_NCELL)
        return STORM_HIP_ERR_INVALID;
    if (c->dense || o.policy == STORM_HIP_POLICY_ALL_BITMAP ||
        o.policy == STORM_HIP_POLICY_EXHAUSTIVE) {
        set_error("thresholded calls take a general corpus and a selecting policy");
        return STORM_HIP_ERR_UNSUPPORTED;
    }
    const uint32_t tile = o.tile ? o.tile : 64;
    SH_TRY(check_tile(tile));
    storm_hip_stats local{};
    storm_hip_stats& s = st ? *st : local;
    std::memset(&s, 0, sizeof s);
    const auto t0 = std::chrono::steady_clock::now();
    SH_HIP(hipSetDevice(ctx->device));
    const uint64_t n = c->n_rows;
    s.pairs = n ? n * (n - 1) / 2 : 0;
    c->hits_valid = false;
    c->h_hits.clear();
    if (n < 2) {
        c->hits_valid = true;
        return STORM_HIP_OK;
    }
    if (!c->d_n_hits) SH_HIP(hipMalloc((void**)&c->d_n_hits, 8));
    if (!c->d_hits) SH_TRY(grow_hits(c, 1u << 16));
    unsigned long long nh = 0;
    SH_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    SH_TRY(threshold_pass(ctx, c, o, tile, t, &s, &nh));
    if (nh > c->hit_cap) {
        // The buffer overflowed: grow it to fit and run again.
        SH_TRY(grow_hits(c, nh + nh / 8));
        SH_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
        SH_TRY(threshold_pass(ctx, c, o, tile, t, &s, &nh));
        if (nh > c->hit_cap) {
            set_error("thresholded call: pair count moved between passes");
            return STORM_HIP_ERR_DEVICE;
        }
    }
    SH_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    std::vector<PairHit> h(nh);
    if (nh) SH_HIP(hipMemcpyAsync(h.data(), c->d_hits, nh * sizeof(PairHit),
                                  hipMemcpyDeviceToHost, ctx->stream));
    SH_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0;
    hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    s.ms_device = ms;
    // Reported in (i, j) order whatever order the device appended them in.
    std::sort(h.begin(), h.end(), [](const PairHit& a, const PairHit& b) {
        return a.i != b.i ? a.i < b.i : a.j < b.j;
    });
    c->h_hits.resize(nh * 3);
    for (uint64_t k = 0; k < nh; ++k) {
        c->h_hits[3 * k] = h[k].i;
        c->h_hits[3 * k + 1] = h[k].j;
        c->h_hits[3 * k + 2] = h[k].count;
    }
    c->hits_valid = true;
    *n_pairs = nh;
    s.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return STORM_HIP_OK;
}

int storm_hip_allpairs_band(storm_hip_ctx* ctx, storm_hip_corpus* c, const storm_hip_options* opt,
                            uint32_t order, uint32_t band, uint32_t* out, storm_hip_stats* st) {
    if (!ctx || !c || c->ctx != ctx || order > STORM_HIP_ORDER_IDENTITY) return STORM_HIP_ERR_INVALID;
    storm_hip_stats local{};
    storm_hip_stats& s = st ? *st : local;
    std::memset(&s, 0, sizeof s);
    const auto t0 = std::chrono::steady_clock::now();
    SH_HIP(hipSetDevice(ctx->device));
    storm_hip_options o;
    uint32_t tile = 0;
    SH_TRY(range_call_options(ctx, c, opt, &o, &tile));
    const uint64_t n = c->n_rows;
    const uint64_t T = (n + tile - 1) / tile;
    if (band >= T) {
        set_error("band %u of a corpus with %llu tile rows", band, (unsigned long long)T);
        return STORM_HIP_ERR_RANGE;
    }
    const uint64_t i0 = (uint64_t)band * tile, i1 = std::min<uint64_t>(n, i0 + tile);
    if (!out) return STORM_HIP_ERR_NULL_VALUES;
    // Pairs (i, j), i in the band, j > i.
    s.pairs = (i1 - i0) * (n - i0) - (i1 - i0) * (i1 - i0 + 1) / 2;
    if (n < 2) {
        std::memset(out, 0, (size_t)((i1 - i0) * n * 4));
        return STORM_HIP_OK;
    }
    SH_TRY(ensure_emit(c, (uint64_t)tile * n * 4));
    Thresh th;
    th.kind = 1;
    th.buf = c->d_emit;
    th.buf_bytes = (uint64_t)tile * n * 4;
    th.identity = order == STORM_HIP_ORDER_IDENTITY;
    th.inv = th.identity ? c->d_ident : c->d_inv;
    const uint32_t sel[4] = {band, band + 1, 0, 0xFFFFFFFFu};
    SH_TRY(run_range(ctx, c, o, tile, th, sel, &s));
    SH_HIP(hipMemcpyAsync(out, c->d_emit, (size_t)((i1 - i0) * n * 4), hipMemcpyDeviceToHost,
                          ctx->stream));
    SH_HIP(hipStreamSynchronize(ctx->stream));
    s.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return STORM_HIP_OK;
}

int storm_hip_allpairs_cross(storm_hip_ctx* ctx, storm_hip_corpus* c, const storm_hip_options* opt,
                             uint32_t split, uint32_t* out, storm_hip_stats* st) {
    if (!ctx || !c || c->ctx != ctx) return STORM_HIP_ERR_INVALID;
    storm_hip_stats local{};
    storm_hip_stats& s = st ? *st : local;
    std::memset(&s, 0, sizeof s);
    const auto t0 = std::chrono::steady_clock::now();
    SH_HIP(hipSetDevice(ctx->device));
    storm_hip_options o;
    uint32_t tile = 0;
    SH_TRY(range_call_options(ctx, c, opt, &o, &tile));
    const uint64_t n = c->n_rows;
    if (split == 0 || split >= n || split % tile != 0) {
        set_error("cross split %u must be a positive multiple of the tile (%u) below the %llu rows",
                  split, tile, (unsigned long long)n);
        return STORM_HIP_ERR_INVALID;
    }
    const uint64_t cols = n - split;
    s.pairs = (uint64_t)split * cols;
    Thresh th;
    th.identity = true;
    if (out) {
        SH_TRY(ensure_emit(c, (uint64_t)split * cols * 4));
        th.kind = 2;
        th.buf = c->d_emit;
        th.buf_bytes = (uint64_t)split * cols * 4;
        th.cols = (uint32_t)cols;
        th.split = split;
        th.inv = c->d_ident;
    }
    const uint32_t sel[4] = {0, split / tile, split / tile, 0xFFFFFFFFu};
    SH_TRY(run_range(ctx, c, o, tile, th, sel, &s));
    if (out) {
        SH_HIP(hipMemcpyAsync(out, c->d_emit, (size_t)split * cols * 4, hipMemcpyDeviceToHost,
                              ctx->stream));
        SH_HIP(hipStreamSynchronize(ctx->stream));
    }
    s.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return STORM_HIP_OK;
}

int storm_hip_corpus_order(storm_hip_ctx* ctx, const storm_hip_corpus* c, uint32_t* perm) {
    if (!ctx || !c || c->ctx != ctx) return STORM_HIP_ERR_INVALID;
    if (!perm && c->n_rows) return STORM_HIP_ERR_NULL_VALUES;
    if (c->dense || c->h_perm.size() != c->n_rows) {
        for (uint32_t k = 0; k < c->n_rows; ++k) perm[k] = k;
    } else {
        std::memcpy(perm, c->h_perm.data(), (size_t)c->n_rows * 4);
    }
    return STORM_HIP_OK;
}

int storm_hip_result_copy_pairs(storm_hip_ctx* ctx, const storm_hip_corpus* c, uint64_t offset,
                                uint64_t count, storm_hip_pair* out) {
    if (!ctx || !c || c->ctx != ctx || !c->hits_valid) return STORM_HIP_ERR_INVALID;
    if (!out && count) return STORM_HIP_ERR_NULL_VALUES;
    const uint64_t total = c->h_hits.size() / 3;
    if (offset > total || count > total - offset) return STORM_HIP_ERR_RANGE;
    for (uint64_t k = 0; k < count; ++k) {
        out[k].i = c->h_hits[3 * (offset + k)];
        out[k].j = c->h_hits[3 * (offset + k) + 1];
        out[k].count = c->h_hits[3 * (offset + k) + 2];
    }
    return STORM_HIP_OK;
}

}  // extern "C"
