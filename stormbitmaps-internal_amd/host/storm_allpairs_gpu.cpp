// storm::gpu::allpairs_sum / allpairs_tiles over the C ABI: the reference's
// batched driver interface (kernels/storm_allpairs.h:196-214) with every
// pair computed on the GPU engine.
#include "storm_allpairs_gpu.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

namespace storm {
namespace gpu {

namespace {

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// The engine's selection tile for a reference tile: 128 and 256 as given,
// anything else 64 (the callback granularity stays the caller's).
uint32_t engine_tile(uint32_t tile) { return (tile == 128 || tile == 256) ? tile : 64; }

// Policy -> options. Oracle makes the same per-pair model choice as PerPair
// (storm_allpairs.cpp:645-660 precomputes it off the clock).
int options_for(Policy policy, uint32_t tile, Pairing fixed_cell, storm_hip_options* o) {
    *o = storm_hip_options{};
    o->tile = engine_tile(tile);
    o->bb_kernel = STORM_HIP_BB_AUTO;
    switch (policy) {
        case Policy::AllBitmap: o->policy = STORM_HIP_POLICY_ALL_BITMAP; break;
        case Policy::PerTile: o->policy = STORM_HIP_POLICY_ANALYTIC; break;
        case Policy::Fixed:
            o->policy = STORM_HIP_POLICY_FIXED;
            o->fixed_cell = (uint32_t)fixed_cell;
            break;
        case Policy::Probe: o->policy = STORM_HIP_POLICY_PROBE; break;
        case Policy::Refine: o->policy = STORM_HIP_POLICY_REFINE; break;
        case Policy::PerPair:
        case Policy::Oracle: o->policy = STORM_HIP_POLICY_PER_PAIR; break;
        default: return STORM_HIP_ERR_UNSUPPORTED;
    }
    return STORM_HIP_OK;
}

// Swaps the context's cost model for one call and restores it.
class ModelScope {
public:
    ModelScope(storm_hip_ctx* ctx, const storm_hip_cost_model* model, bool no_zonemap)
        : ctx_(ctx) {
        if (!model && !no_zonemap) return;
        status_ = storm_hip_get_cost_model(ctx, &saved_);
        if (status_ != STORM_HIP_OK) return;
        storm_hip_cost_model m = model ? *model : saved_;
        if (no_zonemap) m.bb_zonemap = 0.0;
        status_ = storm_hip_set_cost_model(ctx, &m);
        swapped_ = status_ == STORM_HIP_OK;
    }
    ~ModelScope() {
        if (swapped_) (void)storm_hip_set_cost_model(ctx_, &saved_);
    }
    int status() const { return status_; }

private:
    storm_hip_ctx* ctx_;
    storm_hip_cost_model saved_{};
    bool swapped_ = false;
    int status_ = STORM_HIP_OK;
};

void add_stats(const storm_hip_stats& s, AllPairsStats* st) {
    st->sum += s.sum;
    st->pairs += s.pairs;
    st->decisions += s.decisions;
    st->skipped += s.skipped;
    st->ns_selection += s.ms_select * 1e6;
    for (int k = 0; k < (int)Pairing::COUNT && k < STORM_HIP_NCELL; ++k)
        st->cell_pairs[k] += s.cell_pairs[k];
}

int upload_csr(Engine& eng, const RowsCSR& rows, storm_hip_corpus** c) {
    return storm_hip_corpus_from_lists(eng.ctx(), rows.offs, rows.pos, rows.n_rows, rows.universe,
                                       STORM_HIP_BITMAP_DENSE | STORM_HIP_BUILD_RUNS |
                                           STORM_HIP_BUILD_EWAH | STORM_HIP_BUILD_COMP,
                                       c);
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

Engine& default_engine() {
    static Engine eng(0);
    return eng;
}

int infer_universe(RowArenas& a, uint32_t* universe) {
    *universe = 0;
    const uint32_t n = a.n_rows, nw = a.n_words;
    if (nw == 0 || (uint64_t)nw * 64 > 0xFFFFFFFFull) return STORM_HIP_ERR_INVALID;
    uint64_t lo = (uint64_t)(nw - 1) * 64 + 1, hi = (uint64_t)nw * 64, exact = 0;
    // EWAH and complements stay only when every row that should carry one does.
    bool ewah_all = true;
    for (uint32_t r = 0; r < n; ++r) {
        const uint32_t* m = &a.meta[(size_t)r * 8];
        const uint64_t card = m[0];
        const bool bitmap = a.bitmap_off[r + 1] > a.bitmap_off[r];
        if (m[2] != nw) return STORM_HIP_ERR_INVALID;
        if (card) lo = std::max<uint64_t>(lo, (uint64_t)m[5] + 1);
        if (a.ewah_off[r + 1] == a.ewah_off[r]) ewah_all = false;
        if (!bitmap) continue;
        // Zone map iff card * 512 > U.
        if (a.occ_off[r + 1] > a.occ_off[r]) hi = std::min<uint64_t>(hi, card * 512 - 1);
        else lo = std::max<uint64_t>(lo, card * 512);
        // Complement iff 2 * card > U, of exactly U - card positions.
        const uint64_t cl = a.comp_off[r + 1] - a.comp_off[r];
        if (cl) {
            if (exact && exact != card + cl) return STORM_HIP_ERR_INVALID;
            exact = card + cl;
        } else {
            lo = std::max<uint64_t>(lo, 2 * card);
        }
    }
    if (exact) lo = std::max(lo, exact);
    if (lo > hi || (exact && exact != lo)) return STORM_HIP_ERR_INVALID;
    const uint64_t U = lo;
    // A row without a bitmap (build_row's sparse-only rows) holds no
    // complement; when such a row would need one, no row's is used.
    for (uint32_t r = 0; r < n && a.has_comp; ++r) {
        const uint32_t* m = &a.meta[(size_t)r * 8];
        if (a.bitmap_off[r + 1] == a.bitmap_off[r] && 2ull * m[0] > U) a.has_comp = false;
    }
    a.has_ewah = ewah_all;
    *universe = (uint32_t)U;
    return STORM_HIP_OK;
}

int upload_arenas(Engine& eng, RowArenas& a, storm_hip_corpus** out) {
    *out = nullptr;
    if (eng.status() != STORM_HIP_OK) return eng.status();
    uint32_t U = 0;
    const int rc = infer_universe(a, &U);
    if (rc != STORM_HIP_OK) return rc;
    storm_hip_arena ar;
    std::memset(&ar, 0, sizeof ar);
    ar.n_rows = a.n_rows;
    ar.universe = U;
    ar.list_off = a.list_off.data(); ar.list = a.list.data();
    ar.run_off = a.run_off.data(); ar.run_start = a.run_start.data(); ar.run_end = a.run_end.data();
    if (a.has_ewah) { ar.ewah_off = a.ewah_off.data(); ar.ewah = a.ewah.data(); }
    ar.bitmap_off = a.bitmap_off.data(); ar.bitmap = a.bitmap.data();
    ar.rank_off = a.rank_off.data(); ar.rank = a.rank.data();
    ar.occ_off = a.occ_off.data(); ar.occ = a.occ.data();
    if (a.has_comp) { ar.comp_off = a.comp_off.data(); ar.comp = a.comp.data(); }
    ar.meta = a.meta.data();
    return storm_hip_corpus_upload(eng.ctx(), &ar, out);
}

storm_hip_cost_model device_model_for(Engine& eng, const double* ns_per_unit, int n_cells) {
    storm_hip_cost_model m{};
    if (eng.status() != STORM_HIP_OK || storm_hip_get_cost_model(eng.ctx(), &m) != STORM_HIP_OK)
        return m;
    // B x B is select_pairing's seed and keeps its rate.
    for (int k = 1; k < n_cells && k < 10; ++k)
        if (!(ns_per_unit[k] > 0.0)) m.ns_per_unit[k] = 0.0;
    return m;
}

AllPairsStats allpairs_sum(Engine& eng, storm_hip_corpus* c, const storm_hip_cost_model* model,
                           Policy policy, uint32_t tile, bool no_zonemap, Pairing fixed_cell) {
    AllPairsStats st;
    if (eng.status() != STORM_HIP_OK) { st.status = eng.status(); return st; }
    if (!c) { st.status = STORM_HIP_ERR_INVALID; return st; }
    if (storm_hip_corpus_rows(c) < 2) return st;
    const uint64_t t0 = now_ns();
    storm_hip_options o;
    st.status = options_for(policy, tile, fixed_cell, &o);
    if (st.status != STORM_HIP_OK) return st;
    o.out_mode = STORM_HIP_OUT_SUM;
    ModelScope scope(eng.ctx(), model, no_zonemap);
    if ((st.status = scope.status()) != STORM_HIP_OK) return st;
    storm_hip_stats s{};
    st.status = storm_hip_allpairs(eng.ctx(), c, &o, &s);
    if (st.status == STORM_HIP_OK) add_stats(s, &st);
    st.ns_total = (double)(now_ns() - t0);
    return st;
}

AllPairsStats allpairs_tiles(Engine& eng, storm_hip_corpus* c, const storm_hip_cost_model* model,
                             Policy policy, TileCallback cb, void* ctx, uint32_t tile) {
    AllPairsStats st;
    if (eng.status() != STORM_HIP_OK) { st.status = eng.status(); return st; }
    if (!c) { st.status = STORM_HIP_ERR_INVALID; return st; }
    const uint64_t n = storm_hip_corpus_rows(c);
    if (n < 2 || !cb) return st;
    if (tile == 0) { st.status = STORM_HIP_ERR_INVALID; return st; }
    const uint64_t t0 = now_ns();
    storm_hip_options o;
    st.status = options_for(policy, tile, Pairing::BS, &o);
    if (st.status != STORM_HIP_OK) return st;
    // The reference orders by density under PerTile only (storm_allpairs.cpp:
    // 689-690); its other policies run in row order.
    const uint32_t order =
        policy == Policy::PerTile ? STORM_HIP_ORDER_DENSITY : STORM_HIP_ORDER_IDENTITY;
    ModelScope scope(eng.ctx(), model, false);
    if ((st.status = scope.status()) != STORM_HIP_OK) return st;
    const uint64_t E = o.tile;
    // Engine bands held on the host: rows [b*E, b*E + E) x n, by band index.
    std::vector<std::pair<uint64_t, std::vector<uint32_t>>> held;
    auto band_rows = [&](uint64_t b) -> const std::vector<uint32_t>* {
        for (auto& h : held)
            if (h.first == b) return &h.second;
        std::vector<uint32_t> buf((size_t)(std::min<uint64_t>(n, b * E + E) - b * E) * n);
        storm_hip_stats s{};
        st.status = storm_hip_allpairs_band(eng.ctx(), c, &o, order, (uint32_t)b, buf.data(), &s);
        if (st.status != STORM_HIP_OK) return nullptr;
        add_stats(s, &st);
        held.emplace_back(b, std::move(buf));
        return &held.back().second;
    };
    std::vector<uint32_t> counts((size_t)tile * tile);
    std::vector<const uint32_t*> row_ptr(tile);
    for (uint64_t i0 = 0; i0 < n && st.status == STORM_HIP_OK; i0 += tile) {
        const uint64_t i1 = std::min<uint64_t>(i0 + tile, n);
        // Drop bands above this tile row, then point at each of its rows.
        held.erase(std::remove_if(held.begin(), held.end(),
                                  [&](const std::pair<uint64_t, std::vector<uint32_t>>& h) {
                                      return (h.first + 1) * E <= i0;
                                  }),
                   held.end());
        for (uint64_t i = i0; i < i1 && st.status == STORM_HIP_OK; ++i) {
            const std::vector<uint32_t>* b = band_rows(i / E);
            if (b) row_ptr[i - i0] = b->data() + (size_t)(i - (i / E) * E) * n;
        }
        if (st.status != STORM_HIP_OK) break;
        for (uint64_t j0 = i0; j0 < n; j0 += tile) {
            const uint64_t j1 = std::min<uint64_t>(j0 + tile, n);
            std::fill(counts.begin(), counts.end(), 0u);
            for (uint64_t i = i0; i < i1; ++i) {
                const uint32_t* r = row_ptr[i - i0];
                for (uint64_t j = std::max(j0, i + 1); j < j1; ++j)
                    counts[(size_t)(i - i0) * tile + (j - j0)] = r[j];
            }
            cb((uint32_t)i0, (uint32_t)j0, (uint32_t)(i1 - i0), (uint32_t)(j1 - j0),
               counts.data(), ctx);
        }
    }
    st.ns_total = (double)(now_ns() - t0);
    return st;
}

AllPairsStats allpairs_sum(Engine& eng, const RowsCSR& rows, const storm_hip_cost_model* model,
                           Policy policy, uint32_t tile, bool no_zonemap, Pairing fixed_cell) {
    AllPairsStats st;
    if (rows.n_rows < 2) return st;
    const uint64_t t0 = now_ns();
    storm_hip_options o;
    if ((st.status = options_for(policy, tile, fixed_cell, &o)) != STORM_HIP_OK) return st;
    if (eng.status() != STORM_HIP_OK) { st.status = eng.status(); return st; }
    storm_hip_corpus* c = nullptr;
    if ((st.status = upload_csr(eng, rows, &c)) != STORM_HIP_OK) return st;
    st = allpairs_sum(eng, c, model, policy, tile, no_zonemap, fixed_cell);
    storm_hip_corpus_free(c);
    st.ns_total = (double)(now_ns() - t0);
    return st;
}

AllPairsStats allpairs_tiles(Engine& eng, const RowsCSR& rows, const storm_hip_cost_model* model,
                             Policy policy, TileCallback cb, void* ctx, uint32_t tile) {
    AllPairsStats st;
    if (rows.n_rows < 2 || !cb) return st;
    if (tile == 0) { st.status = STORM_HIP_ERR_INVALID; return st; }
    const uint64_t t0 = now_ns();
    storm_hip_options o;
    if ((st.status = options_for(policy, tile, Pairing::BS, &o)) != STORM_HIP_OK) return st;
    if (eng.status() != STORM_HIP_OK) { st.status = eng.status(); return st; }
    storm_hip_corpus* c = nullptr;
    if ((st.status = upload_csr(eng, rows, &c)) != STORM_HIP_OK) return st;
    st = allpairs_tiles(eng, c, model, policy, cb, ctx, tile);
    storm_hip_corpus_free(c);
    st.ns_total = (double)(now_ns() - t0);
    return st;
}

}  // namespace gpu
}  // namespace storm
