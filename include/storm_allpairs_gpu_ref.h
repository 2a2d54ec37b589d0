/* storm::gpu overloads with the reference's exact parameter lists
 * (kernels/storm_allpairs.h:196-214), for code built against the reference's
 * headers: include this after kernels/storm_allpairs.h and call
 * storm::gpu::allpairs_sum / allpairs_tiles where storm::allpairs_sum /
 * allpairs_tiles were called (or bring them in with a using-declaration).
 *
 * The rows' arenas are uploaded as built by build_row (list, runs, EWAH,
 * bitmap, rank9, zone map, complement and RowMeta; storm_repr.h:237-276) to
 * the process-wide default engine (device 0) and checked there against each
 * row. Results (sum, pairs, per-pair counts) equal the reference's; stats
 * count the engine's own decisions. A failed GPU call returns empty stats
 * (pairs 0) and leaves the status in storm::gpu::last_status(), the message
 * in storm_hip_last_error().
 *
 * CostModel: the reference's rates are CPU nanoseconds, so the device prices
 * cells with its own rates; a cell the caller's model leaves unpriced
 * (ns_per_unit <= 0) is excluded on the device too. */
#ifndef STORM_ALLPAIRS_GPU_REF_H_
#define STORM_ALLPAIRS_GPU_REF_H_

#include <vector>

#include "storm_allpairs_gpu.h"

namespace storm {
namespace gpu {

/* Status of the calling thread's last reference-typed call. */
inline int& last_status() {
    static thread_local int s = STORM_HIP_OK;
    return s;
}

namespace ref_detail {

template <class V, class T>
inline void append(std::vector<T>& dst, std::vector<uint64_t>& off, const V& src) {
    dst.insert(dst.end(), src.data(), src.data() + src.size());
    off.push_back(dst.size());
}

template <class RowT>
inline int arenas_of(const std::vector<RowT>& rows, RowArenas& a) {
    a.n_rows = (uint32_t)rows.size();
    a.n_words = rows.empty() ? 0 : rows[0].meta.n_words;
    for (auto* off : {&a.list_off, &a.run_off, &a.ewah_off, &a.bitmap_off, &a.rank_off,
                      &a.occ_off, &a.comp_off})
        off->assign(1, 0);
    a.meta.reserve((size_t)rows.size() * 8);
    for (const RowT& r : rows) {
        // Zone maps of 8 words per bin are the only layout the device reads.
        if (!r.occ.empty() && r.occ_bin != 8) return STORM_HIP_ERR_UNSUPPORTED;
        append(a.list, a.list_off, r.list);
        a.run_start.insert(a.run_start.end(), r.run_start.data(),
                           r.run_start.data() + r.run_start.size());
        append(a.run_end, a.run_off, r.run_end);
        append(a.ewah, a.ewah_off, r.ewah);
        append(a.bitmap, a.bitmap_off, r.bitmap);
        append(a.rank, a.rank_off, r.rank);
        append(a.occ, a.occ_off, r.occ);
        append(a.comp, a.comp_off, r.comp_list);
        const uint32_t m[8] = {r.meta.cardinality, r.meta.n_runs,    r.meta.n_words,
                               r.meta.n_nonzero_w, r.meta.first_set, r.meta.last_set,
                               (uint32_t)r.meta.has_rank, (uint32_t)r.meta.has_occ};
        a.meta.insert(a.meta.end(), m, m + 8);
    }
    return STORM_HIP_OK;
}

template <class StatsT>
inline StatsT to_ref(const AllPairsStats& s) {
    StatsT r;
    last_status() = s.status;
    if (s.status != STORM_HIP_OK) return r;
    r.sum = s.sum;
    r.pairs = s.pairs;
    r.decisions = s.decisions;
    r.ns_total = s.ns_total;
    r.ns_selection = s.ns_selection;
    r.skipped = s.skipped;
    const size_t nc = sizeof(r.cell_pairs) / sizeof(r.cell_pairs[0]);
    for (size_t k = 0; k < nc && k < (size_t)Pairing::COUNT; ++k) r.cell_pairs[k] = s.cell_pairs[k];
    return r;
}

template <class RowT, class ModelT>
inline int upload(const std::vector<RowT>& rows, const ModelT& model, storm_hip_corpus** c,
                  storm_hip_cost_model* dm) {
    Engine& eng = default_engine();
    if (eng.status() != STORM_HIP_OK) return eng.status();
    RowArenas a;
    int rc = arenas_of(rows, a);
    if (rc == STORM_HIP_OK) rc = upload_arenas(eng, a, c);
    if (rc == STORM_HIP_OK)
        *dm = device_model_for(eng, model.ns_per_unit,
                               (int)(sizeof(model.ns_per_unit) / sizeof(model.ns_per_unit[0])));
    return rc;
}

}  // namespace ref_detail

/* kernels/storm_allpairs.h:196-201 */
inline ::storm::AllPairsStats allpairs_sum(const std::vector<::storm::Row>& rows,
                                           const ::storm::CostModel& model,
                                           ::storm::Policy policy, uint32_t tile = 64,
                                           bool no_zonemap = false,
                                           ::storm::Pairing fixed_cell = ::storm::Pairing::BS) {
    if (rows.size() < 2) { last_status() = STORM_HIP_OK; return ::storm::AllPairsStats{}; }
    storm_hip_corpus* c = nullptr;
    storm_hip_cost_model dm{};
    AllPairsStats s;
    s.status = ref_detail::upload(rows, model, &c, &dm);
    if (s.status == STORM_HIP_OK)
        s = allpairs_sum(default_engine(), c, &dm, (Policy)(uint8_t)policy, tile, no_zonemap,
                         (Pairing)(uint8_t)fixed_cell);
    if (c) storm_hip_corpus_free(c);
    return ref_detail::to_ref<::storm::AllPairsStats>(s);
}

/* kernels/storm_allpairs.h:206-214 */
inline ::storm::AllPairsStats allpairs_tiles(const std::vector<::storm::Row>& rows,
                                             const ::storm::CostModel& model,
                                             ::storm::Policy policy, ::storm::TileCallback cb,
                                             void* ctx, uint32_t tile = 64) {
    if (rows.size() < 2 || !cb) { last_status() = STORM_HIP_OK; return ::storm::AllPairsStats{}; }
    storm_hip_corpus* c = nullptr;
    storm_hip_cost_model dm{};
    AllPairsStats s;
    s.status = ref_detail::upload(rows, model, &c, &dm);
    if (s.status == STORM_HIP_OK)
        s = allpairs_tiles(default_engine(), c, &dm, (Policy)(uint8_t)policy,
                           reinterpret_cast<TileCallback>(cb), ctx, tile);
    if (c) storm_hip_corpus_free(c);
    return ref_detail::to_ref<::storm::AllPairsStats>(s);
}

}  // namespace gpu
}  // namespace storm

#endif
