// storm::gpu::allpairs_sum / allpairs_tiles against brute force, in the shape
// of the reference's check_allpairs (tests/test_cells.cpp:237-288): every
// policy, both API forms, tiles 8 / 16 / 64 / 100 / 128 (several tiles per
// engine band, bands per callback tile), several corpus structures.
// Needs a GPU; prints one line per failure and exits non-zero on any.
#include <algorithm>
#include <cinttypes>
#include <cstdio>
#include <set>
#include <vector>

#include "storm_allpairs_gpu.h"

using namespace storm::gpu;

namespace {

uint64_t g_checks = 0, g_fails = 0;

void check(bool ok, const char* what, const char* ctx, uint64_t got, uint64_t want) {
    ++g_checks;
    if (!ok) {
        ++g_fails;
        std::printf("FAIL %s [%s]: got %" PRIu64 " want %" PRIu64 "\n", what, ctx, got, want);
    }
}

struct Xorshift {   // xorshift64*
    uint64_t s;
    uint64_t next() {
        s ^= s >> 12; s ^= s << 25; s ^= s >> 27;
        return s * 0x2545F4914F6CDD1Dull;
    }
    uint32_t below(uint32_t n) { return (uint32_t)(next() % n); }
};

// structure 0: uniform positions; 1: clustered (a window per row); 2: runs.
// spectrum 0: equal sizes; 1: sizes falling as 1/rank.
std::vector<std::vector<uint32_t>> corpus(uint32_t n, uint32_t u, double d, int structure,
                                          int spectrum, uint64_t seed) {
    Xorshift g{seed | 1};
    std::vector<std::vector<uint32_t>> rows(n);
    for (uint32_t r = 0; r < n; ++r) {
        uint32_t k = (uint32_t)(d * u * (spectrum ? 3.0 / (1 + r % 7) : 1.0));
        k = std::min(k, u);
        std::set<uint32_t> s;
        if (structure == 0) {
            while (s.size() < k) s.insert(g.below(u));
        } else if (structure == 1) {
            const uint32_t w = std::min(u, std::max(k * 4, 64u)), lo = g.below(u - w + 1);
            while (s.size() < k) s.insert(lo + g.below(w));
        } else {
            uint32_t p = g.below(u);
            while (s.size() < k) {
                const uint32_t len = 1 + g.below(40);
                for (uint32_t q = p; q < std::min(u, p + len) && s.size() < k; ++q) s.insert(q);
                p = (p + len + g.below(200)) % u;
            }
        }
        rows[r].assign(s.begin(), s.end());
    }
    if (n > 1) rows[1].clear();   // an empty row
    return rows;
}

uint64_t inter(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {
    uint64_t c = 0;
    size_t i = 0, j = 0;
    while (i < a.size() && j < b.size()) {
        if (a[i] < b[j]) ++i;
        else if (b[j] < a[i]) ++j;
        else { ++c; ++i; ++j; }
    }
    return c;
}

// The reference's coordinates: density-order positions under PerTile (stable
// by cardinality, kernels/storm_allpairs.cpp:316-323), row order otherwise.
std::vector<uint32_t> order_for(Policy pol, const std::vector<std::vector<uint32_t>>& rows) {
    std::vector<uint32_t> o(rows.size());
    for (uint32_t k = 0; k < o.size(); ++k) o[k] = k;
    if (pol == Policy::PerTile)
        std::stable_sort(o.begin(), o.end(), [&](uint32_t a, uint32_t b) {
            return rows[a].size() < rows[b].size();
        });
    return o;
}

struct CbState {
    const std::vector<uint32_t>* full;   // n x n brute-force counts
    uint32_t n;
    std::vector<uint32_t> ord;
    uint32_t tile;
    uint64_t sum = 0, bad = 0, calls = 0;
};

void on_tile(uint32_t i0, uint32_t j0, uint32_t nr, uint32_t nc, const uint32_t* cnt, void* vp) {
    CbState* s = static_cast<CbState*>(vp);
    ++s->calls;
    for (uint32_t a = 0; a < nr; ++a)
        for (uint32_t b = 0; b < nc; ++b) {
            const uint32_t i = i0 + a, j = j0 + b;
            if (j <= i) continue;
            const uint32_t v = cnt[(size_t)a * s->tile + b];
            s->sum += v;
            if (v != (*s->full)[(size_t)s->ord[i] * s->n + s->ord[j]]) ++s->bad;
        }
}

}  // namespace

int main() {
    Engine eng(0);
    if (eng.status() != STORM_HIP_OK) {
        std::printf("no device: %s\n", storm_hip_last_error());
        return 2;
    }
    for (uint32_t u : {1024u, 4033u, 16384u})
        for (int st = 0; st < 3; ++st)
            for (int sp = 0; sp < 2; ++sp) {
                const auto rows = corpus(300, u, 0.02, st, sp, 0x9911u + u * 7 + st * 3 + sp);
                std::vector<uint64_t> offs{0};
                std::vector<uint32_t> pos;
                for (const auto& r : rows) {
                    pos.insert(pos.end(), r.begin(), r.end());
                    offs.push_back(pos.size());
                }
                RowsCSR R;
                R.offs = offs.data(); R.pos = pos.data(); R.n_rows = (uint32_t)rows.size();
                R.universe = u;
                uint64_t want = 0;
                const size_t n = rows.size();
                std::vector<uint32_t> full(n * n, 0);
                for (size_t i = 0; i < n; ++i)
                    for (size_t j = i + 1; j < n; ++j) {
                        full[i * n + j] = full[j * n + i] = (uint32_t)inter(rows[i], rows[j]);
                        want += full[i * n + j];
                    }
                char ctx[96];
                for (Policy pol : {Policy::AllBitmap, Policy::PerTile, Policy::Probe, Policy::Fixed}) {
                    for (uint32_t tile : {8u, 16u, 64u, 100u, 128u}) {
                        std::snprintf(ctx, sizeof ctx, "%s u=%u st=%d sp=%d tile=%u", name_of(pol),
                                      u, st, sp, tile);
                        const AllPairsStats s1 = allpairs_sum(eng, R, nullptr, pol, tile, false,
                                                              Pairing::SS);
                        check(s1.status == STORM_HIP_OK, "status", ctx, (uint64_t)-s1.status, 0);
                        check(s1.sum == want, "allpairs_sum", ctx, s1.sum, want);
                        check(s1.pairs == 300ull * 299 / 2, "pairs", ctx, s1.pairs, 300ull * 299 / 2);
                        CbState cs;
                        cs.full = &full;
                        cs.n = (uint32_t)n;
                        cs.ord = order_for(pol, rows);
                        cs.tile = tile;
                        const AllPairsStats s2 = allpairs_tiles(eng, R, nullptr, pol, on_tile, &cs, tile);
                        check(s2.status == STORM_HIP_OK, "tiles status", ctx, (uint64_t)-s2.status, 0);
                        check(s2.sum == want, "allpairs_tiles", ctx, s2.sum, want);
                        check(cs.sum == want, "callback sum", ctx, cs.sum, want);
                        check(cs.bad == 0, "callback per-pair", ctx, cs.bad, 0);
                        const uint64_t T = (300 + tile - 1) / tile;
                        check(cs.calls == T * (T + 1) / 2, "callback tiles", ctx, cs.calls, T * (T + 1) / 2);
                    }
                }
                // The per-pair policies.
                for (Policy pol : {Policy::PerPair, Policy::Oracle, Policy::Refine}) {
                    const AllPairsStats sp_ = allpairs_sum(eng, R, nullptr, pol);
                    check(sp_.status == STORM_HIP_OK && sp_.sum == want, "per-pair policies",
                          name_of(pol), sp_.sum, want);
                }
            }
    // Fewer than two rows: an empty result, as the reference returns.
    RowsCSR one;
    const uint64_t o1[2] = {0, 0};
    one.offs = o1; one.n_rows = 1; one.universe = 64;
    const AllPairsStats s0 = allpairs_sum(eng, one, nullptr, Policy::PerTile);
    check(s0.sum == 0 && s0.pairs == 0 && s0.status == STORM_HIP_OK, "one row", "", s0.pairs, 0);
    std::printf("%" PRIu64 " checks, %" PRIu64 " failures\n", g_checks, g_fails);
    return g_fails ? 1 : 0;
}
