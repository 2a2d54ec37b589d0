// STORM_t on the GPU: the host keeps every vector as a sorted position list;
// the pairwise calls upload them (once per change) as a general corpus and run
// the engine's per-tile selection on the device. Results equal those of
// storm.cpp:872-1046 on sorted input.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "storm_gpu.h"
#include "storm_hip.h"

namespace {

constexpr uint32_t kBlockValues = 65536;      // STORM_DEFAULT_BLOCK_SIZE
constexpr uint32_t kScalarThreshold = 4096;   // STORM_DEFAULT_SCALAR_THRESHOLD

// Serialized bytes of one vector in the reference's layout (storm.cpp:396-418):
// per block, u16 scalars (fewer than 4,096 values) or 1,024 words, plus a
// 16-byte header; per vector, 4 bytes per block id plus 12.
uint64_t serialized_bytes(const uint32_t* v, uint32_t n) {
    uint64_t total = 12;
    uint32_t start = 0;
    while (start < n) {
        uint32_t stop = start;
        const uint32_t block = v[start] / kBlockValues;
        while (stop < n && v[stop] / kBlockValues == block) ++stop;
        const uint32_t c = stop - start;
        total += 16 + 4 + (c < kScalarThreshold ? 2ull * c : 1024ull * 8);
        start = stop;
    }
    return total;
}

// Generations are unique over the process (a freed handle's address may be
// reused by a new one, its generation never).
uint64_t next_gen() {
    static std::atomic<uint64_t> g{0};
    return ++g;
}

}  // namespace

struct STORM_s {
    std::vector<std::vector<uint32_t>> rows;   // sorted, distinct
    uint64_t ser_bytes = 8;                    // STORM_serialized_size
    uint64_t gen = next_gen();                 // renewed by every add / clear
    mutable storm_hip_ctx* ctx = nullptr;
    mutable storm_hip_corpus* corpus = nullptr;
    mutable bool dirty = true;
    // The last square call with this handle first: the two-set corpus, the
    // other handle and both generations it was built from.
    mutable storm_hip_corpus* sq_corpus = nullptr;
    mutable const STORM_s* sq_other = nullptr;
    mutable uint64_t sq_gen_self = 0, sq_gen_other = 0;
    mutable uint32_t sq_split = 0;
};

namespace {

void report(int s, const char* what) {
    std::fprintf(stderr, "%s: GPU call failed (%s): %s\n", what, storm_hip_strerror(s),
                 storm_hip_last_error());
}

int ensure_ctx(const STORM_t* h) {
    if (h->ctx) return STORM_HIP_OK;
    return storm_hip_ctx_create(0, 0, &h->ctx);
}

// Sum over pairs within rows[0..n) of |Ri n Rj| on the device.
int pairwise_sum(storm_hip_ctx* ctx, const std::vector<const std::vector<uint32_t>*>& rows,
                 storm_hip_corpus** cache, uint64_t* out) {
    *out = 0;
    if (rows.size() < 2) return STORM_HIP_OK;
    storm_hip_corpus* c = cache ? *cache : nullptr;
    if (!c) {
        std::vector<uint64_t> offs{0};
        std::vector<uint32_t> pos;
        uint64_t universe = 1;
        for (const auto* r : rows) {
            pos.insert(pos.end(), r->begin(), r->end());
            offs.push_back(pos.size());
            if (!r->empty()) universe = std::max<uint64_t>(universe, (uint64_t)r->back() + 1);
        }
        const int s = storm_hip_corpus_from_lists(
            ctx, offs.data(), pos.data(), (uint32_t)rows.size(), (uint32_t)universe,
            STORM_HIP_BITMAP_DENSE | STORM_HIP_BUILD_RUNS | STORM_HIP_BUILD_EWAH |
                STORM_HIP_BUILD_COMP,
            &c);
        if (s != STORM_HIP_OK) return s;
        if (cache) *cache = c;
    }
    storm_hip_options o;
    std::memset(&o, 0, sizeof o);
    o.policy = STORM_HIP_POLICY_ANALYTIC;
    o.out_mode = STORM_HIP_OUT_SUM;
    o.tile = 64;
    storm_hip_stats st;
    const int s = storm_hip_allpairs(ctx, c, &o, &st);
    if (!cache) storm_hip_corpus_free(c);
    if (s != STORM_HIP_OK) return s;
    *out = st.sum;
    return STORM_HIP_OK;
}

// Corpus of the rows of a then b, a padded with empty rows to a multiple
// of kCrossTile; each side in density order (stable by cardinality), which
// keeps tiles homogeneous and does not change the sum over the block.
constexpr uint32_t kCrossTile = 64;

int cross_corpus(storm_hip_ctx* ctx, const STORM_s* a, const STORM_s* b, storm_hip_corpus** out,
                 uint32_t* split) {
    std::vector<uint64_t> offs{0};
    std::vector<uint32_t> pos;
    uint64_t universe = 1;
    auto add_side = [&](const STORM_s* h) {
        std::vector<uint32_t> ord(h->rows.size());
        for (uint32_t k = 0; k < ord.size(); ++k) ord[k] = k;
        std::stable_sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) {
            return h->rows[x].size() < h->rows[y].size();
        });
        for (uint32_t k : ord) {
            const auto& r = h->rows[k];
            pos.insert(pos.end(), r.begin(), r.end());
            offs.push_back(pos.size());
            if (!r.empty()) universe = std::max<uint64_t>(universe, (uint64_t)r.back() + 1);
        }
    };
    add_side(a);
    const uint64_t n1 = a->rows.size();
    *split = (uint32_t)((n1 + kCrossTile - 1) / kCrossTile * kCrossTile);
    for (uint64_t k = n1; k < *split; ++k) offs.push_back(pos.size());
    add_side(b);
    if (offs.size() - 1 > 0xFFFFFFFFull) return STORM_HIP_ERR_RANGE;
    return storm_hip_corpus_from_lists(ctx, offs.data(), pos.data(), (uint32_t)(offs.size() - 1),
                                       (uint32_t)universe,
                                       STORM_HIP_BITMAP_DENSE | STORM_HIP_BUILD_RUNS |
                                           STORM_HIP_BUILD_EWAH | STORM_HIP_BUILD_COMP,
                                       out);
}

uint64_t pairw(STORM_t* h, const char* what) {
    if (h == nullptr) return UINT64_MAX;
    if (h->rows.size() < 2) return 0;
    int s = ensure_ctx(h);
    if (s == STORM_HIP_OK && h->dirty && h->corpus) {
        storm_hip_corpus_free(h->corpus);
        h->corpus = nullptr;
    }
    uint64_t sum = 0;
    if (s == STORM_HIP_OK) {
        std::vector<const std::vector<uint32_t>*> rows;
        for (const auto& r : h->rows) rows.push_back(&r);
        s = pairwise_sum(h->ctx, rows, &h->corpus, &sum);
        if (s == STORM_HIP_OK) h->dirty = false;
    }
    if (s != STORM_HIP_OK) {
        report(s, what);
        return UINT64_MAX;
    }
    return sum;
}

}  // namespace

extern "C" {

STORM_t* STORM_new(void) { return new (std::nothrow) STORM_s(); }

void STORM_free(STORM_t* bitmap) {
    if (bitmap == nullptr) return;
    if (bitmap->corpus) storm_hip_corpus_free(bitmap->corpus);
    if (bitmap->sq_corpus) storm_hip_corpus_free(bitmap->sq_corpus);
    if (bitmap->ctx) storm_hip_ctx_free(bitmap->ctx);
    delete bitmap;
}

int STORM_add(STORM_t* bitmap, const uint32_t* values, const uint32_t n_values) {
    if (bitmap == nullptr) return -1;
    std::vector<uint32_t> v;
    if (values && n_values) {
        if (std::find(values, values + n_values, 0xFFFFFFFFu) != values + n_values) return -3;
        try {
            v.assign(values, values + n_values);
        } catch (...) {
            return -4;
        }
        bitmap->ser_bytes += serialized_bytes(values, n_values);
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    } else {
        bitmap->ser_bytes += 12;
    }
    try {
        bitmap->rows.push_back(std::move(v));
    } catch (...) {
        return -4;
    }
    bitmap->dirty = true;
    bitmap->gen = next_gen();
    return 1;
}

int STORM_clear(STORM_t* bitmap) {
    if (bitmap == nullptr) return -1;
    bitmap->rows.clear();
    bitmap->ser_bytes = 8;
    bitmap->dirty = true;
    bitmap->gen = next_gen();
    return 1;
}

uint64_t STORM_pairw_intersect_cardinality(STORM_t* bitmap) {
    return pairw(bitmap, "STORM_pairw_intersect_cardinality");
}

uint64_t STORM_pairw_intersect_cardinality_blocked(STORM_t* bitmap, uint32_t bsize) {
    (void)bsize;
    return pairw(bitmap, "STORM_pairw_intersect_cardinality_blocked");
}

// Every (i in the first, j in the second) pair and nothing else: one corpus
// holding both sets, and only the tile pairs across the split dispatched
// (n1 x n2 pair evaluations, as storm.cpp:1027-1046 visits). The corpus is
// kept on the first handle until either handle changes.
uint64_t STORM_intersect_cardinality_square(const STORM_t* bitmap1, const STORM_t* bitmap2) {
    if (bitmap1 == nullptr || bitmap2 == nullptr) return 0;
    if (bitmap1->rows.empty() || bitmap2->rows.empty()) return 0;
    int s = ensure_ctx(bitmap1);
    const bool cached = bitmap1->sq_corpus && bitmap1->sq_other == bitmap2 &&
                        bitmap1->sq_gen_self == bitmap1->gen &&
                        bitmap1->sq_gen_other == bitmap2->gen;
    if (s == STORM_HIP_OK && !cached) {
        if (bitmap1->sq_corpus) storm_hip_corpus_free(bitmap1->sq_corpus);
        bitmap1->sq_corpus = nullptr;
        s = cross_corpus(bitmap1->ctx, bitmap1, bitmap2, &bitmap1->sq_corpus, &bitmap1->sq_split);
        if (s == STORM_HIP_OK) {
            bitmap1->sq_other = bitmap2;
            bitmap1->sq_gen_self = bitmap1->gen;
            bitmap1->sq_gen_other = bitmap2->gen;
        }
    }
    storm_hip_stats st{};
    if (s == STORM_HIP_OK) {
        storm_hip_options o;
        std::memset(&o, 0, sizeof o);
        o.policy = STORM_HIP_POLICY_ANALYTIC;
        o.out_mode = STORM_HIP_OUT_SUM;
        o.tile = kCrossTile;
        s = storm_hip_allpairs_cross(bitmap1->ctx, bitmap1->sq_corpus, &o, bitmap1->sq_split,
                                     nullptr, &st);
    }
    if (s != STORM_HIP_OK) {
        report(s, "STORM_intersect_cardinality_square");
        return 0;
    }
    return st.sum;
}

uint64_t STORM_serialized_size(const STORM_t* bitmap) {
    if (bitmap == nullptr) return 0;
    return bitmap->ser_bytes;
}

}  // extern "C"
