/*
 * C entry points over the compiled StormBitmaps reference, for tests and for
 * the CPU baseline. TEST INFRASTRUCTURE ONLY: built by oracle/Makefile from
 * the reference sources where they lie, into oracle/_ref/, never shipped.
 *
 * Each function names the reference routine it drives.
 */
#include "storm.h"
#include "kernels/storm_allpairs.h"
#include "kernels/storm_cells.h"
#include "kernels/storm_cost.h"
#include "kernels/storm_gen.h"
#include "kernels/storm_repr.h"

#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

using namespace storm;

extern "C" {

/* First n draws of storm::Rng(seed).next() (kernels/storm_gen.h:84-102). */
void ref_rng_next(uint64_t seed, uint32_t n, uint64_t* out) {
    Rng r(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = r.next();
}

/* n draws of storm::Rng(seed).below(bound) (kernels/storm_gen.h:93-99). */
void ref_rng_below(uint64_t seed, uint32_t bound, uint32_t n, uint32_t* out) {
    Rng r(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = r.below(bound);
}

/* build_row (kernels/storm_repr.cpp:102-207): sizes first, then contents.
 * sizes[0..8] = n_words, n_rank, n_occ, n_runs, n_ewah, n_comp, n_comp_runs,
 * has_rank, has_occ. Arrays may be NULL to query sizes only. meta6 receives
 * cardinality, n_runs, n_words, n_nonzero_w, first_set, last_set. */
int ref_build_row(const uint32_t* pos, uint32_t n, uint32_t universe, uint32_t* sizes,
                  uint32_t* meta6, uint64_t* bitmap, uint64_t* rank, uint64_t* occ,
                  uint32_t* run_start, uint32_t* run_end, uint64_t* ewah,
                  uint32_t* comp_list) {
    Row r;
    build_row(r, pos, n, universe);
    sizes[0] = r.meta.n_words; sizes[1] = (uint32_t)r.rank.size();
    sizes[2] = (uint32_t)r.occ.size(); sizes[3] = (uint32_t)r.run_start.size();
    sizes[4] = (uint32_t)r.ewah.size(); sizes[5] = (uint32_t)r.comp_list.size();
    sizes[6] = (uint32_t)r.comp_start.size();
    sizes[7] = r.meta.has_rank; sizes[8] = r.meta.has_occ;
    meta6[0] = r.meta.cardinality; meta6[1] = r.meta.n_runs; meta6[2] = r.meta.n_words;
    meta6[3] = r.meta.n_nonzero_w; meta6[4] = r.meta.first_set; meta6[5] = r.meta.last_set;
    if (bitmap) std::memcpy(bitmap, r.bitmap.data(), r.bitmap.size() * 8);
    if (rank && !r.rank.empty()) std::memcpy(rank, r.rank.data(), r.rank.size() * 8);
    if (occ && !r.occ.empty()) std::memcpy(occ, r.occ.data(), r.occ.size() * 8);
    if (run_start) std::memcpy(run_start, r.run_start.data(), r.run_start.size() * 4);
    if (run_end) std::memcpy(run_end, r.run_end.data(), r.run_end.size() * 4);
    if (ewah) std::memcpy(ewah, r.ewah.data(), r.ewah.size() * 8);
    if (comp_list && !r.comp_list.empty())
        std::memcpy(comp_list, r.comp_list.data(), r.comp_list.size() * 4);
    return 0;
}

/* build_rank (kernels/storm_repr.cpp:25-45), unconditionally. */
void ref_build_rank(const uint64_t* w, uint32_t nw, uint64_t* out) {
    avec<uint64_t> r;
    build_rank(w, nw, r);
    std::memcpy(out, r.data(), r.size() * 8);
}

static RowMeta meta_from(const uint32_t* m) {
    RowMeta r;
    r.cardinality = m[0]; r.n_runs = m[1]; r.n_words = m[2]; r.n_nonzero_w = m[3];
    r.first_set = m[4]; r.last_set = m[5]; r.has_rank = m[6] != 0; r.has_occ = m[7] != 0;
    return r;
}

/* select_pairing + predict with default_model (kernels/storm_cost.cpp:372-496).
 * meta arrays are 8 u32: the six RowMeta integers then has_rank, has_occ. */
int ref_select_pairing(const uint32_t* a, const uint32_t* b, double* predicted) {
    CostModel m; default_model(m);
    return (int)select_pairing(m, meta_from(a), meta_from(b), predicted);
}

double ref_predict(int p, const uint32_t* a, const uint32_t* b) {
    CostModel m; default_model(m);
    return predict(m, (Pairing)p, meta_from(a), meta_from(b));
}

static void rows_from_csr(const uint32_t* offs, const uint32_t* pos, uint32_t n,
                          uint32_t universe, std::vector<Row>& rows) {
    rows.resize(n);
    for (uint32_t i = 0; i < n; ++i)
        build_row(rows[i], pos + offs[i], offs[i + 1] - offs[i], universe);
}

/* allpairs_sum (kernels/storm_allpairs.cpp:486-675) over CSR rows with the
 * default model. policy: 0 AllBitmap, 1 PerPair, 2 PerTile, 3 Oracle, 4 Fixed,
 * 5 Probe, 6 Refine. cell_pairs receives 11 counters. Returns the sum. */
uint64_t ref_allpairs_sum(const uint32_t* offs, const uint32_t* pos, uint32_t n,
                          uint32_t universe, int policy, uint32_t tile,
                          uint64_t* cell_pairs, uint64_t* pairs, uint64_t* skipped) {
    std::vector<Row> rows;
    rows_from_csr(offs, pos, n, universe, rows);
    CostModel m; default_model(m);
    AllPairsStats s = allpairs_sum(rows, m, (Policy)policy, tile);
    if (cell_pairs) for (int k = 0; k < (int)Pairing::COUNT; ++k) cell_pairs[k] = s.cell_pairs[k];
    if (pairs) *pairs = s.pairs;
    if (skipped) *skipped = s.skipped;
    return s.sum;
}

/* allpairs_tiles (kernels/storm_allpairs.cpp:677-732) with identity order
 * (AllBitmap): scatter tile counts into a packed upper triangle. */
struct TriCtx { uint32_t* out; uint32_t n; uint32_t tile; };
static void tri_cb(uint32_t i0, uint32_t j0, uint32_t nr, uint32_t nc,
                   const uint32_t* counts, void* vp) {
    TriCtx* c = (TriCtx*)vp;
    for (uint32_t a = 0; a < nr; ++a)
        for (uint32_t b = 0; b < nc; ++b) {
            const uint64_t i = i0 + a, j = j0 + b;
            if (j <= i) continue;
            c->out[i * c->n - i * (i + 1) / 2 + (j - i - 1)] = counts[(size_t)a * c->tile + b];
        }
}
uint64_t ref_allpairs_tiles_tri(const uint32_t* offs, const uint32_t* pos, uint32_t n,
                                uint32_t universe, uint32_t tile, uint32_t* out_tri) {
    std::vector<Row> rows;
    rows_from_csr(offs, pos, n, universe, rows);
    CostModel m; default_model(m);
    TriCtx c{out_tri, n, tile};
    return allpairs_tiles(rows, m, Policy::AllBitmap, tri_cb, &c, tile).sum;
}

/* generate (kernels/storm_gen.cpp:173-232) into CSR: returns total elements;
 * call with pos == NULL first to size. structure 0/1/2, spectrum 0/1/2. */
uint64_t ref_generate(uint32_t n_rows, uint32_t universe, double density, int structure,
                      int spectrum, uint64_t seed, uint32_t* offs, uint32_t* pos) {
    CorpusSpec spec;
    spec.n_rows = n_rows; spec.universe = universe; spec.density = density;
    spec.structure = (Structure)structure; spec.spectrum = (Spectrum)spectrum; spec.seed = seed;
    Corpus c; generate(c, spec);
    uint64_t tot = 0;
    for (uint32_t i = 0; i < n_rows; ++i) {
        if (offs) offs[i] = (uint32_t)tot;
        if (pos) std::memcpy(pos + tot, c.rows[i].list.data(), c.rows[i].list.size() * 4);
        tot += c.rows[i].list.size();
    }
    if (offs) offs[n_rows] = (uint32_t)tot;
    return tot;
}

/* CPU baseline leg (i), the reference as shipped on one thread: rows built
 * with build_row (kernels/storm_repr.cpp:102-207, untimed), then
 * allpairs_sum (kernels/storm_allpairs.cpp:486-675) under `policy`
 * (0 AllBitmap, 2 PerTile) with the default model or, when calibrate_first,
 * the model calibrate() (kernels/storm_cost.cpp:556-769) fits on this host
 * (untimed). Returns the sum; *seconds is allpairs_sum's wall time. */
uint64_t ref_time_allpairs(const uint32_t* offs, const uint32_t* pos, uint32_t n,
                           uint32_t universe, int policy, uint32_t tile, int calibrate_first,
                           double* seconds) {
    std::vector<Row> rows;
    rows_from_csr(offs, pos, n, universe, rows);
    CostModel m; default_model(m);
    if (calibrate_first) calibrate(m);
    auto t0 = std::chrono::steady_clock::now();
    AllPairsStats s = allpairs_sum(rows, m, (Policy)policy, tile);
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return s.sum;
}

/* CPU baseline leg (i) for the contiguous container: n rows of nw words (row
 * major) added to a STORM_contiguous_t of n_bits (storm.cpp:1159-1220,
 * untimed), then STORM_contig_pairw_intersect_cardinality (storm.cpp:1232-1240)
 * timed on one thread. */
uint64_t ref_time_contig(const uint64_t* words, uint32_t n, uint32_t nw, uint32_t n_bits,
                         double* seconds) {
    STORM_contiguous_t* c = STORM_contig_new(n_bits);
    if (!c) return 0;
    std::vector<uint32_t> p;
    for (uint32_t i = 0; i < n; ++i) {
        p.clear();
        for (uint32_t w = 0; w < nw; ++w) {
            uint64_t x = words[(size_t)i * nw + w];
            while (x) { p.push_back(w * 64 + (uint32_t)__builtin_ctzll(x)); x &= x - 1; }
        }
        STORM_contig_add(c, p.data(), (uint32_t)p.size());
    }
    auto t0 = std::chrono::steady_clock::now();
    const uint64_t s = STORM_contig_pairw_intersect_cardinality(c);
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    STORM_contig_free(c);
    return s;
}

/* CPU baseline: the reference's shipped all-bitmap B x B kernel ("dense",
 * kernels/cell_bb.cpp:402-404 -> storm_simd.h:144-203) over all pairs of the
 * given rows, with project threading: row i of the upper triangle goes to
 * worker i % threads. Returns the sum; *seconds receives wall time. */
uint64_t ref_bb_dense_allpairs(const uint64_t* words, uint32_t n, uint32_t nw, int threads,
                               double* seconds) {
    const VariantList<fn_bb> vl = cell_bb();
    fn_bb dense = vl.v[0].fn;
    for (size_t i = 0; i < vl.n; ++i)
        if (std::strcmp(vl.v[i].name, "dense") == 0) dense = vl.v[i].fn;
    if (threads < 1) threads = 1;
    std::vector<uint64_t> sums((size_t)threads, 0);
    auto t0 = std::chrono::steady_clock::now();
    auto work = [&](int t) {
        uint64_t s = 0;
        for (uint32_t i = (uint32_t)t; i < n; i += (uint32_t)threads) {
            BitmapView a{words + (size_t)i * nw, nw, nullptr, nullptr, 0, 8};
            for (uint32_t j = i + 1; j < n; ++j) {
                BitmapView b{words + (size_t)j * nw, nw, nullptr, nullptr, 0, 8};
                s += dense(a, b);
            }
        }
        sums[(size_t)t] = s;
    };
    std::vector<std::thread> th;
    for (int t = 1; t < threads; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    uint64_t s = 0;
    for (uint64_t v : sums) s += v;
    return s;
}

} // extern "C"
