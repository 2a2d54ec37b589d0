// Seeded synthetic corpora for the benchmark workloads (host side).
//
// The reference generator (kernels/storm_gen.{h,cpp}) has no Bernoulli, Zipf or
// Poisson spectrum, so these are new; they reuse its splitmix64 stream
// definition (kernels/storm_gen.h:84-102: next, below, unit). Every row draws
// from its OWN stream, seeded by draw r of Rng(seed), so rows generate in
// parallel and any single row can be regenerated alone. Integer sampling only
// where the result must not depend on libm (positions); the Zipf inverse CDF
// and the geometric/Poisson transforms use the deterministic log/exp of
// detmath.h on the same draws, so the device generators (gen_dev.hip) write
// the same rows.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "detmath.h"
#include "device_common.h"
#include "internal.h"

namespace stormhip {
namespace {

using Rng = DetRng;

inline uint64_t row_seed(uint64_t seed, uint64_t r) { return det_row_seed(seed, r); }

// k distinct uniform positions in [0, u), sorted: draw, sort, unique, top up.
void distinct_uniform(Rng& g, uint32_t u, uint32_t k, std::vector<uint32_t>& out) {
    out.clear();
    if (k >= u) { out.resize(u); for (uint32_t i = 0; i < u; ++i) out[i] = i; return; }
    if ((uint64_t)k * 2 > u) {   // sample the complement
        std::vector<uint32_t> drop;
        distinct_uniform(g, u, u - k, drop);
        out.reserve(k);
        size_t d = 0;
        for (uint32_t i = 0; i < u; ++i) {
            if (d < drop.size() && drop[d] == i) { ++d; continue; }
            out.push_back(i);
        }
        return;
    }
    while (out.size() < k) {
        const size_t need = k - out.size();
        for (size_t q = 0; q < need; ++q) out.push_back(g.below(u));
        std::sort(out.begin(), out.end());
        out.erase(std::unique(out.begin(), out.end()), out.end());
    }
}

// Runs + singletons row (det_runs_row), as sorted distinct positions.
void run_row(Rng& g, uint32_t u, double lambda, double mean_run, double frac,
             std::vector<uint32_t>& out) {
    out.clear();
    det_runs_row(g, u, lambda, mean_run, frac, [&](uint32_t s, uint32_t e) {
        for (uint32_t p = s; p < e; ++p) out.push_back(p);
    });
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
}

// Bernoulli(d) row by geometric skips: exact i.i.d. bits, O(card).
void bernoulli_row(Rng& g, uint32_t u, double d, std::vector<uint32_t>& out) {
    out.clear();
    if (d <= 0.0) return;
    if (d >= 1.0) { out.resize(u); for (uint32_t i = 0; i < u; ++i) out[i] = i; return; }
    const double l1 = det_log1p(-d);
    for (double pos = det_bernoulli_next(g, l1, -1.0); pos < (double)u;
         pos = det_bernoulli_next(g, l1, pos))
        out.push_back((uint32_t)pos);
}

template <class F>
int gen_rows(uint32_t n, F&& row_fn, uint64_t** offs_out, uint32_t** pos_out, uint64_t* total) {
    std::vector<std::vector<uint32_t>> rows(n);
    const unsigned th = std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
    std::atomic<uint32_t> next{0};
    auto work = [&] {
        for (uint32_t r; (r = next.fetch_add(1)) < n;) row_fn(r, rows[r]);
    };
    std::vector<std::thread> ts;
    for (unsigned t = 1; t < th; ++t) ts.emplace_back(work);
    work();
    for (auto& t : ts) t.join();
    uint64_t tot = 0;
    for (auto& r : rows) tot += r.size();
    uint64_t* offs = (uint64_t*)std::malloc(sizeof(uint64_t) * ((size_t)n + 1));
    uint32_t* pos = (uint32_t*)std::malloc(sizeof(uint32_t) * (tot ? tot : 1));
    if (!offs || !pos) { std::free(offs); std::free(pos); return STORM_HIP_ERR_NOMEM; }
    offs[0] = 0;
    for (uint32_t r = 0; r < n; ++r) {
        offs[r + 1] = offs[r] + rows[r].size();
        if (!rows[r].empty()) std::memcpy(pos + offs[r], rows[r].data(), rows[r].size() * 4);
        std::vector<uint32_t>().swap(rows[r]);
    }
    *offs_out = offs; *pos_out = pos; *total = tot;
    return STORM_HIP_OK;
}

}  // namespace
}  // namespace stormhip

using namespace stormhip;

extern "C" {

int storm_hip_gen_corpus(uint32_t kind, uint32_t n_rows, uint32_t universe, double p1,
                         double p2, double p3, uint64_t seed, uint64_t** offs, uint32_t** pos,
                         uint64_t* total) {
    if (!offs || !pos || !total || universe == 0) return STORM_HIP_ERR_INVALID;
    switch (kind) {
        case STORM_HIP_GEN_ZIPF: {
            // Cardinality Zipf(s = p1) on [1, K = p2] by inverse CDF, positions
            // uniform without replacement.
            const uint32_t K = (uint32_t)std::min<double>(p2, universe);
            if (K == 0 || p1 <= 0) return STORM_HIP_ERR_INVALID;
            std::vector<double> cdf(K);
            double acc = 0;
            for (uint32_t k = 1; k <= K; ++k) { acc += std::pow((double)k, -p1); cdf[k - 1] = acc; }
            for (auto& x : cdf) x /= acc;
            return gen_rows(n_rows, [&](uint32_t r, std::vector<uint32_t>& out) {
                Rng g(row_seed(seed, r));
                const double u = g.unit();
                const uint32_t k = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
                distinct_uniform(g, universe, std::min(k, K), out);
            }, offs, pos, total);
        }
        case STORM_HIP_GEN_RUNS:
            // Poisson(p1) runs of geometric length (mean p2) + p3 singletons.
            return gen_rows(n_rows, [&](uint32_t r, std::vector<uint32_t>& out) {
                Rng g(row_seed(seed, r));
                run_row(g, universe, p1, p2, p3, out);
            }, offs, pos, total);
        case STORM_HIP_GEN_BERNOULLI:
            return gen_rows(n_rows, [&](uint32_t r, std::vector<uint32_t>& out) {
                Rng g(row_seed(seed, r));
                bernoulli_row(g, universe, p1, out);
            }, offs, pos, total);
        default:
            return STORM_HIP_ERR_INVALID;
    }
}

void storm_hip_gen_free(void* p) { std::free(p); }

}  // extern "C"
