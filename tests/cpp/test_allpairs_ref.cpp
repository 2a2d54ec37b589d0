// The reference-typed storm::gpu overloads (storm_allpairs_gpu_ref.h) against
// the reference's own storm::allpairs_sum / allpairs_tiles on the same
// std::vector<Row> (kernels/storm_allpairs.h:196-214): equal sums and pair
// counts, and the very same sequence of tile callbacks (coordinates, shapes
// and counts) under every policy both run. Built against the reference's
// headers and sources (oracle/Makefile); needs a GPU. Prints one line per
// failure and exits non-zero on any.
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <vector>

#include "kernels/storm_allpairs.h"
#include "kernels/storm_cost.h"
#include "kernels/storm_gen.h"
#include "kernels/storm_repr.h"
#include "storm_allpairs_gpu_ref.h"

namespace {

uint64_t g_checks = 0, g_fails = 0;

void check(bool ok, const char* what, const char* ctx, uint64_t got, uint64_t want) {
    ++g_checks;
    if (!ok) {
        ++g_fails;
        std::printf("FAIL %s [%s]: got %" PRIu64 " want %" PRIu64 "\n", what, ctx, got, want);
    }
}

// Every callback as (i0, j0, nr, nc, counts of the nr x nc window).
struct Trace {
    uint32_t tile = 0;
    std::vector<uint32_t> log;
};

void record(uint32_t i0, uint32_t j0, uint32_t nr, uint32_t nc, const uint32_t* counts, void* vp) {
    Trace* t = static_cast<Trace*>(vp);
    t->log.insert(t->log.end(), {i0, j0, nr, nc});
    for (uint32_t a = 0; a < nr; ++a)
        t->log.insert(t->log.end(), counts + (size_t)a * t->tile, counts + (size_t)a * t->tile + nc);
}

std::vector<storm::Row> rows_of(const std::vector<std::vector<uint32_t>>& lists, uint32_t u,
                                uint32_t sparse_every) {
    std::vector<storm::Row> rows(lists.size());
    for (size_t r = 0; r < lists.size(); ++r)
        storm::build_row(rows[r], lists[r].data(), lists[r].size(), u,
                         sparse_every && r % sparse_every == 1);
    return rows;
}

std::vector<std::vector<uint32_t>> lists_of(const std::vector<storm::Row>& rows) {
    std::vector<std::vector<uint32_t>> out;
    for (const auto& r : rows) out.emplace_back(r.list.begin(), r.list.end());
    return out;
}

void run_case(const char* name, const std::vector<storm::Row>& rows, const storm::CostModel& m) {
    using storm::Policy;
    char ctx[160];
    for (Policy pol : {Policy::AllBitmap, Policy::PerTile, Policy::Fixed, Policy::Probe,
                       Policy::Refine, Policy::PerPair, Policy::Oracle}) {
        for (uint32_t tile : {8u, 64u, 100u, 256u}) {
            std::snprintf(ctx, sizeof ctx, "%s %s tile=%u", name, storm::name_of(pol), tile);
            std::printf("case %s\n", ctx);
            std::fflush(stdout);
            const storm::AllPairsStats want = storm::allpairs_sum(rows, m, pol, tile, false,
                                                                  storm::Pairing::SS);
            const storm::AllPairsStats got = storm::gpu::allpairs_sum(rows, m, pol, tile, false,
                                                                      storm::Pairing::SS);
            check(storm::gpu::last_status() == STORM_HIP_OK, "status", ctx,
                  (uint64_t)-storm::gpu::last_status(), 0);
            check(got.sum == want.sum, "allpairs_sum", ctx, got.sum, want.sum);
            check(got.pairs == want.pairs, "pairs", ctx, got.pairs, want.pairs);
            Trace tw, tg;
            tw.tile = tg.tile = tile;
            const storm::AllPairsStats sw = storm::allpairs_tiles(rows, m, pol, record, &tw, tile);
            const storm::AllPairsStats sg = storm::gpu::allpairs_tiles(rows, m, pol, record, &tg, tile);
            check(storm::gpu::last_status() == STORM_HIP_OK, "tiles status", ctx,
                  (uint64_t)-storm::gpu::last_status(), 0);
            check(sg.sum == sw.sum, "allpairs_tiles sum", ctx, sg.sum, sw.sum);
            check(tg.log.size() == tw.log.size(), "callback trace length", ctx, tg.log.size(),
                  tw.log.size());
            size_t diff = 0;
            for (size_t k = 0; k < tg.log.size() && k < tw.log.size(); ++k) diff += tg.log[k] != tw.log[k];
            check(diff == 0, "callback trace", ctx, diff, 0);
        }
    }
    // no_zonemap and another forced cell, sums only.
    for (storm::Pairing fc : {storm::Pairing::BB, storm::Pairing::RR}) {
        std::snprintf(ctx, sizeof ctx, "%s fixed=%d no_zonemap", name, (int)fc);
        const auto want = storm::allpairs_sum(rows, m, Policy::Fixed, 64, true, fc);
        const auto got = storm::gpu::allpairs_sum(rows, m, Policy::Fixed, 64, true, fc);
        check(got.sum == want.sum && got.pairs == want.pairs, "fixed no_zonemap", ctx, got.sum,
              want.sum);
    }
}

}  // namespace

int main() {
    if (storm::gpu::default_engine().status() != STORM_HIP_OK) {
        std::printf("no device: %s\n", storm_hip_last_error());
        return 2;
    }
    storm::CostModel dm;
    storm::default_model(dm);
    const storm::CostModel zero;   // uncalibrated: every cell but B x B excluded
    const storm::Structure structures[] = {storm::Structure::Uniform, storm::Structure::Clustered,
                                           storm::Structure::Runs};
    const storm::Spectrum spectra[] = {storm::Spectrum::Uniform, storm::Spectrum::Inverse,
                                       storm::Spectrum::Bimodal};
    int k = 0;
    for (uint32_t u : {4033u, 65536u})
        for (storm::Structure st : structures)
            for (storm::Spectrum sp : spectra) {
                storm::CorpusSpec spec;
                spec.n_rows = 300;
                spec.universe = u;
                spec.density = u > 10000 ? 0.01 : 0.05;
                spec.structure = st;
                spec.spectrum = sp;
                spec.seed = 0xC0FFEEull + (uint64_t)k++;
                storm::Corpus c;
                storm::generate(c, spec);
                char name[96];
                std::snprintf(name, sizeof name, "u=%u %s %s", u, storm::name_of(st),
                              storm::name_of(sp));
                run_case(name, c.rows, k % 3 == 0 ? zero : dm);
            }
    // Rows built sparse-only (no bitmap, EWAH, indexes or complement) beside
    // full rows, and rows dense enough to carry complements.
    {
        storm::CorpusSpec spec;
        spec.n_rows = 260;
        spec.universe = 20000;
        spec.density = 0.02;
        spec.spectrum = storm::Spectrum::Inverse;
        storm::Corpus c;
        storm::generate(c, spec);
        auto lists = lists_of(c.rows);
        for (size_t r = 0; r < lists.size(); r += 9) {   // every 9th row past half density
            lists[r].clear();
            for (uint32_t p = 0; p < spec.universe; ++p)
                if ((p * 2654435761u) % 7 < 5) lists[r].push_back(p);
        }
        run_case("complements", rows_of(lists, spec.universe, 0), dm);
        // The reference runs sparse-only rows only through cells that never
        // touch a bitmap (storm_repr.h:280-289): compare forced SS / RR sums,
        // and the GPU's tile walk against the same total.
        const auto sparse = rows_of(lists_of(c.rows), spec.universe, 4);
        for (storm::Pairing fc : {storm::Pairing::SS, storm::Pairing::RR}) {
            const auto want = storm::allpairs_sum(sparse, dm, storm::Policy::Fixed, 64, false, fc);
            for (storm::Policy pol : {storm::Policy::Fixed, storm::Policy::PerTile,
                                      storm::Policy::AllBitmap}) {
                const auto got = storm::gpu::allpairs_sum(sparse, dm, pol, 64, false, fc);
                check(storm::gpu::last_status() == STORM_HIP_OK && got.sum == want.sum &&
                          got.pairs == want.pairs,
                      "sparse-only rows", storm::name_of(pol), got.sum, want.sum);
                Trace tg;
                tg.tile = 64;
                const auto sg = storm::gpu::allpairs_tiles(sparse, dm, pol, record, &tg, 64);
                check(storm::gpu::last_status() == STORM_HIP_OK && sg.sum == want.sum,
                      "sparse-only rows tiles", storm::name_of(pol), sg.sum, want.sum);
            }
        }
    }
    // Fewer than two rows: empty stats, as the reference.
    {
        std::vector<storm::Row> one(1);
        const uint32_t p = 5;
        storm::build_row(one[0], &p, 1, 64);
        const auto got = storm::gpu::allpairs_sum(one, dm, storm::Policy::PerTile);
        check(got.pairs == 0 && got.sum == 0, "one row", "", got.pairs, 0);
    }
    std::printf("%" PRIu64 " checks, %" PRIu64 " failures\n", g_checks, g_fails);
    return g_fails ? 1 : 0;
}
