// The sparse and compressed cells: exact |Xi n Xj| for every pair of a tile of
// rows, one kernel per cell class, persistent workgroups pulling tiles from
// that class's device queue.
//
// Cell semantics follow the reference's cell families (kernels/cell_bs.cpp,
// cell_br.cpp, cell_sparse.cpp, cell_wah.cpp, cell_comp.cpp) and the driver's
// orientation rules (kernels/storm_allpairs.cpp:192-251):
//   B x S  probe the bitmap at each list position
//   B x R  masked popcount of the bitmap over each run
//   B x W  walk the EWAH markers: one-fills popcount the bitmap, literals AND
//   S x S  each element of the shorter list searched in the longer one
//   S x R  each list element located among the runs (the sparser row's list,
//          the denser row's runs)
//   S x W  each list element's word looked up in the EWAH marker table
//   R x R  overlap of each run of the side with fewer runs with the other
//   R x W  bits of the EWAH row inside each run
//   W x W  markers of one row against the other's decoded words
//   C x B  |A n B| = |B| - |A' n B| over the complement list A' (cell_comp.cpp)
// Unlike the reference driver, S x W and R x W are dispatched (it returns 0
// for them, storm_allpairs.cpp:245-249); every count here is exact.
//
// Work split: a pair's work items (list elements, runs, markers) are spread
// over the 64 lanes of a wave and wave-reduced ("wave mode"), or, for tiles
// whose rows are tiny, each lane evaluates whole pairs ("lane mode").
// Three cells also work on whole blocks of pairs at once, touching each row's
// items once per block instead of once per pair: S x S as an LDS hash join and
// R x R as an LDS interval join over 64 x 64 blocks, and B x S as stripes of
// 64 list rows against a 256-row bitmap tile read through its bit slices.
#include <atomic>

#include "device_common.h"
#include "general.h"
#include "cost_model.h"

namespace stormhip {

namespace {

struct Span32 { const uint32_t* p; uint32_t n; };
struct RunsV { const uint32_t* s; const uint32_t* e; uint32_t n; };
struct EwahV { const uint64_t* buf; const uint32_t* mw; const uint32_t* mp; uint32_t nm; };

__device__ inline Span32 list_of(const DevCorpus& D, uint32_t r) {
    const uint64_t a = D.list_off[r];
    return {D.list + a, (uint32_t)(D.list_off[r + 1] - a)};
}
__device__ inline Span32 comp_of(const DevCorpus& D, uint32_t r) {
    const uint64_t a = D.comp_off[r];
    return {D.comp + a, (uint32_t)(D.comp_off[r + 1] - a)};
}
__device__ inline RunsV runs_of(const DevCorpus& D, uint32_t r) {
    const uint64_t a = D.run_off[r];
    return {D.run_start + a, D.run_end + a, (uint32_t)(D.run_off[r + 1] - a)};
}
__device__ inline EwahV ewah_of(const DevCorpus& D, uint32_t r) {
    const uint64_t m = D.mk_off[r];
    return {D.ewah + D.ewah_off[r], D.mk_word + m, D.mk_pos + m, (uint32_t)(D.mk_off[r + 1] - m)};
}
__device__ inline const uint64_t* bitmap_of(const DevCorpus& D, uint32_t r) {
    const uint32_t b = D.bm_row[r];
    return b == kNoBitmap ? nullptr : D.bitmap + (uint64_t)b * D.ld;
}
__device__ inline uint32_t meta(const DevCorpus& D, uint32_t r, int k) { return D.meta[(uint64_t)r * 8 + k]; }
__device__ inline const uint64_t* rank_of(const DevCorpus& D, uint32_t r) {
    const uint32_t k = D.rank_row[r];
    return k == kNoBitmap ? nullptr : D.rank + (uint64_t)k * D.rank_ld;
}
__device__ inline const uint64_t* occ_of(const DevCorpus& D, uint32_t r) {
    const uint32_t k = D.occ_row[r];
    return k == kNoBitmap ? nullptr : D.occ + (uint64_t)k * D.occ_ld;
}

// Set bits strictly below position x through the rank9 index (rank_at,
// kernels/storm_repr.h:129-137): block count, 9-bit in-block count, one word.
__device__ inline uint64_t rank_at(const DevCorpus& D, const uint64_t* bm, const uint64_t* rk,
                                   uint32_t x) {
    const uint32_t wi = x >> 6;
    if (wi >= D.nw) return rk[2 * ((D.nw + 7) / 8)];
    const uint32_t blk = wi >> 3, k = wi & 7;
    uint64_t r = rk[2 * blk];
    if (k) r += (rk[2 * blk + 1] >> (9 * (k - 1))) & 0x1FF;
    return r + __popcll(bm[wi] & ((1ull << (x & 63)) - 1));
}
// Runs at least this many bits long are counted by two rank lookups, shorter
// ones by a masked popcount (br_hybrid<4>, kernels/cell_br.cpp:187-199).
constexpr uint32_t kRankMinBits = 4 * 64;

// First index with a[idx] >= x.
__device__ inline uint32_t lower_bound(const uint32_t* a, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}
// lower_bound for an x expected near the front of a: probes 1, 2, 4, ...
// elements ahead, then bisects the last step, so the reads stay near a[0]
// (the join kernels' chunk boundaries advance a few dozen elements at a time).
__device__ inline uint32_t gallop_lower_bound(const uint32_t* a, uint32_t n, uint32_t x) {
    if (n == 0 || a[0] >= x) return 0;
    uint32_t lo = 0, step = 1;   // a[lo] < x throughout
    while (lo + step < n && a[lo + step] < x) { lo += step; step <<= 1; }
    uint32_t hi = min(n, lo + step);   // a[hi] >= x, or hi == n
    ++lo;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}
// First index with a[idx] > x.
__device__ inline uint32_t upper_bound(const uint32_t* a, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ inline uint64_t bits_from(uint32_t lo) { return ~0ull << lo; }          // bits >= lo
__device__ inline uint64_t bits_to(uint32_t hi) { return ~0ull >> (63 - hi); }     // bits <= hi

// Set bits of a bitmap inside [s, e).
__device__ inline uint32_t bitmap_range_count(const uint64_t* bm, uint32_t s, uint32_t e) {
    if (s >= e) return 0;
    const uint32_t ws = s >> 6, we = (e - 1) >> 6;
    if (ws == we) return __popcll(bm[ws] & bits_from(s & 63) & bits_to((e - 1) & 63));
    uint32_t c = __popcll(bm[ws] & bits_from(s & 63)) + __popcll(bm[we] & bits_to((e - 1) & 63));
    for (uint32_t w = ws + 1; w < we; ++w) c += __popcll(bm[w]);
    return c;
}

__device__ inline uint32_t ewah_find(const EwahV& E, uint32_t w) {
    const uint32_t k = upper_bound(E.mw, E.nm, w);
    return k ? k - 1 : 0;
}
__device__ inline uint64_t ewah_word(const EwahV& E, uint32_t m, uint32_t w) {
    const uint64_t mk = E.buf[E.mp[m]];
    const uint32_t fl = (uint32_t)((mk >> 1) & 0x7FFFFFFFull), ll = (uint32_t)(mk >> 32);
    uint32_t off = w - E.mw[m];
    if (off < fl) return (mk & 1) ? ~0ull : 0ull;
    off -= fl;
    return off < ll ? E.buf[E.mp[m] + 1 + off] : 0ull;
}
__device__ inline uint32_t ewah_next_start(const EwahV& E, uint32_t m) {
    return m + 1 < E.nm ? E.mw[m + 1] : 0xFFFFFFFFu;
}

// Set bits of an EWAH row inside [s, e); fills are counted without a word walk.
__device__ inline uint32_t ewah_range_count(const EwahV& E, uint32_t s, uint32_t e) {
    if (s >= e || E.nm == 0) return 0;
    const uint32_t ws = s >> 6, we = (e - 1) >> 6;
    uint32_t m = ewah_find(E, ws), w = ws, c = 0;
    while (w <= we && m < E.nm) {
        const uint64_t mk = E.buf[E.mp[m]];
        const uint32_t fl = (uint32_t)((mk >> 1) & 0x7FFFFFFFull), ll = (uint32_t)(mk >> 32);
        const uint32_t fill_end = E.mw[m] + fl, lit_end = fill_end + ll;
        if (w < fill_end) {
            const uint32_t hi = min(fill_end - 1, we);
            if (mk & 1) {
                const uint64_t a = max((uint64_t)s, (uint64_t)w * 64);
                const uint64_t b = min((uint64_t)e, ((uint64_t)hi + 1) * 64);
                c += (uint32_t)(b - a);
            }
            w = hi + 1;
            continue;
        }
        if (w < lit_end) {
            uint64_t v = E.buf[E.mp[m] + 1 + (w - fill_end)];
            if (w == ws) v &= bits_from(s & 63);
            if (w == we) v &= bits_to((e - 1) & 63);
            c += __popcll(v);
            ++w;
            continue;
        }
        ++m;
    }
    return c;
}

// One cell evaluated for pair (i, j) (original ids), over work items
// first, first+stride, ... . Returns this thread's partial count.
template <int CELL>
__device__ uint32_t cell_partial(const DevCorpus& D, const DevModel& M, uint32_t i, uint32_t j,
                                 uint32_t first, uint32_t stride) {
    const uint32_t ci = meta(D, i, 0), cj = meta(D, j, 0);
    const bool i_dense = ci >= cj;
    const uint32_t d = i_dense ? i : j, s = i_dense ? j : i;
    uint32_t c = 0;
    if constexpr (CELL == STORM_HIP_CELL_BS) {
        const uint64_t* bm = bitmap_of(D, d);
        uint32_t l = s;
        if (!bm) { bm = bitmap_of(D, s); l = d; }
        const Span32 L = list_of(D, l);
#pragma unroll 4
        for (uint32_t k = first; k < L.n; k += stride) {
            const uint32_t x = L.p[k];
            c += (uint32_t)(bm[x >> 6] >> (x & 63)) & 1u;
        }
    } else if constexpr (CELL == STORM_HIP_CELL_BR || CELL == STORM_HIP_CELL_BW) {
        // Run-count orientation with the reference's margin (storm_cost.cpp:181-188).
        const double rd = meta(D, d, 1), rs = meta(D, s, 1);
        bool swap = M.br_margin > 0.0 && rd * M.br_margin <= rs;
        uint32_t b = swap ? s : d, o = swap ? d : s;
        if (!bitmap_of(D, b)) { const uint32_t t = b; b = o; o = t; }
        const uint64_t* bm = bitmap_of(D, b);
        if constexpr (CELL == STORM_HIP_CELL_BR) {
            const RunsV R = runs_of(D, o);
            const uint64_t* rk = rank_of(D, b);
            if (rk) {
                for (uint32_t k = first; k < R.n; k += stride) {
                    const uint32_t rs = R.s[k], re = R.e[k];
                    c += re - rs >= kRankMinBits
                             ? (uint32_t)(rank_at(D, bm, rk, re) - rank_at(D, bm, rk, rs))
                             : bitmap_range_count(bm, rs, re);
                }
            } else {
#pragma unroll 4
                for (uint32_t k = first; k < R.n; k += stride) c += bitmap_range_count(bm, R.s[k], R.e[k]);
            }
        } else {
            const EwahV E = ewah_of(D, o);
            for (uint32_t k = first; k < E.nm; k += stride) {
                const uint64_t mk = E.buf[E.mp[k]];
                const uint32_t fl = (uint32_t)((mk >> 1) & 0x7FFFFFFFull), ll = (uint32_t)(mk >> 32);
                const uint32_t w0 = E.mw[k];
                if (mk & 1)
                    for (uint32_t w = w0; w < w0 + fl; ++w) c += __popcll(bm[w]);
                for (uint32_t q = 0; q < ll; ++q)
                    c += __popcll(bm[w0 + fl + q] & E.buf[E.mp[k] + 1 + q]);
            }
        }
    } else if constexpr (CELL == STORM_HIP_CELL_SS) {
        const Span32 A = list_of(D, s), B = list_of(D, d);   // |A| <= |B|
#pragma unroll 4
        for (uint32_t k = first; k < A.n; k += stride) {
            const uint32_t x = A.p[k];
            const uint32_t at = lower_bound(B.p, B.n, x);
            c += (at < B.n && B.p[at] == x) ? 1u : 0u;
        }
    } else if constexpr (CELL == STORM_HIP_CELL_SR) {
        const Span32 L = list_of(D, s);
        const RunsV R = runs_of(D, d);
#pragma unroll 4
        for (uint32_t k = first; k < L.n; k += stride) {
            const uint32_t x = L.p[k];
            const uint32_t u = upper_bound(R.s, R.n, x);
            c += (u > 0 && x < R.e[u - 1]) ? 1u : 0u;
        }
    } else if constexpr (CELL == STORM_HIP_CELL_SW) {
        const Span32 L = list_of(D, s);
        const EwahV E = ewah_of(D, d);
        for (uint32_t k = first; k < L.n; k += stride) {
            const uint32_t x = L.p[k];
            const uint64_t v = ewah_word(E, ewah_find(E, x >> 6), x >> 6);
            c += (uint32_t)(v >> (x & 63)) & 1u;
        }
    } else if constexpr (CELL == STORM_HIP_CELL_RR) {
        uint32_t a = d, b = s;
        if (meta(D, s, 1) < meta(D, d, 1)) { a = s; b = d; }
        const RunsV A = runs_of(D, a), B = runs_of(D, b);
        for (uint32_t k = first; k < A.n; k += stride) {
            const uint32_t rs = A.s[k], re = A.e[k];
            for (uint32_t q = upper_bound(B.e, B.n, rs); q < B.n && B.s[q] < re; ++q)
                c += min(re, B.e[q]) - max(rs, B.s[q]);
        }
    } else if constexpr (CELL == STORM_HIP_CELL_RW) {
        // min(ra + wb, rb + wa) orientation (storm_cost.cpp:329).
        const uint64_t ra = meta(D, d, 1), rb = meta(D, s, 1);
        const uint64_t wa = D.ewah_off[d + 1] - D.ewah_off[d], wb = D.ewah_off[s + 1] - D.ewah_off[s];
        const uint32_t r_side = (ra + wb <= rb + wa) ? d : s, w_side = (r_side == d) ? s : d;
        const RunsV R = runs_of(D, r_side);
        const EwahV E = ewah_of(D, w_side);
        for (uint32_t k = first; k < R.n; k += stride) c += ewah_range_count(E, R.s[k], R.e[k]);
    } else if constexpr (CELL == STORM_HIP_CELL_WW) {
        uint32_t a = d, b = s;
        if (D.mk_off[s + 1] - D.mk_off[s] < D.mk_off[d + 1] - D.mk_off[d]) { a = s; b = d; }
        const EwahV A = ewah_of(D, a), B = ewah_of(D, b);
        for (uint32_t k = first; k < A.nm; k += stride) {
            const uint64_t mk = A.buf[A.mp[k]];
            const uint32_t fl = (uint32_t)((mk >> 1) & 0x7FFFFFFFull), ll = (uint32_t)(mk >> 32);
            const uint32_t w0 = A.mw[k];
            if ((mk & 1) && fl) c += ewah_range_count(B, w0 * 64u, (w0 + fl) * 64u);
            if (ll) {
                uint32_t m = ewah_find(B, w0 + fl), next = ewah_next_start(B, m);
                for (uint32_t q = 0; q < ll; ++q) {
                    const uint32_t w = w0 + fl + q;
                    while (w >= next) { ++m; next = ewah_next_start(B, m); }
                    c += __popcll(A.buf[A.mp[k] + 1 + q] & ewah_word(B, m, w));
                }
            }
        }
    } else if constexpr (CELL == STORM_HIP_CELL_CB) {
        // Complement side a (2n > universe), bitmap side b. Partial = probes of
        // the complement into b; the caller forms |b| - probes.
        uint32_t a = d, b = s;
        if (!(2ull * meta(D, d, 0) > D.universe) || !bitmap_of(D, s)) { a = s; b = d; }
        const uint64_t* bm = bitmap_of(D, b);
        const Span32 A = comp_of(D, a);
        const RunsV R = runs_of(D, a);
        // cb_adaptive (kernels/cell_comp.cpp:70-74): the complement as runs
        // (the gaps between the row's runs) when it averages at least 4
        // positions per gap or no complement list is held, else as positions.
        const uint32_t comp_n = D.universe - meta(D, a, 0);
        if (R.n && (A.n != comp_n || comp_n >= 4u * (R.n + 1))) {
            const uint64_t* rk = rank_of(D, b);
            // Gap k: [end of run k-1, start of run k), with run -1 ending at 0
            // and run R.n starting at the universe.
            for (uint32_t k = first; k <= R.n; k += stride) {
                const uint32_t gs = k ? R.e[k - 1] : 0, ge = k < R.n ? R.s[k] : D.universe;
                if (gs >= ge) continue;
                c += rk && ge - gs >= kRankMinBits
                         ? (uint32_t)(rank_at(D, bm, rk, ge) - rank_at(D, bm, rk, gs))
                         : bitmap_range_count(bm, gs, ge);
            }
        } else {
#pragma unroll 4
            for (uint32_t k = first; k < A.n; k += stride) {
                const uint32_t x = A.p[k];
                c += (uint32_t)(bm[x >> 6] >> (x & 63)) & 1u;
            }
        }
    }
    return c;
}

// The card of the bitmap side the C x B partial was taken against.
__device__ inline uint32_t cb_base(const DevCorpus& D, uint32_t i, uint32_t j) {
    const uint32_t ci = meta(D, i, 0), cj = meta(D, j, 0);
    const uint32_t d = ci >= cj ? i : j, s = ci >= cj ? j : i;
    uint32_t b = s;
    if (!(2ull * meta(D, d, 0) > D.universe) || !bitmap_of(D, s)) b = d;
    return meta(D, b, 0);
}

// Span test (kernels/storm_allpairs.cpp:627-634): true if provably disjoint.
// Thresholded calls also skip a pair whose length bound min/max is below the
// threshold (it cannot be reported, whatever its count).
__device__ inline bool span_disjoint(const DevCorpus& D, uint32_t i, uint32_t j) {
    const uint32_t ci = meta(D, i, 0), cj = meta(D, j, 0);
    if (ci == 0 || cj == 0 || meta(D, i, 5) < meta(D, j, 4) || meta(D, j, 5) < meta(D, i, 4))
        return true;
    return D.thr > 0.0 && (double)min(ci, cj) / (double)max(ci, cj) < D.thr;
}
// The same test on two rows' (cardinality, first, last) records.
__device__ inline bool span_disjoint(const DevCorpus& D, uint4 a, uint4 b) {
    if (a.x == 0 || b.x == 0 || a.z < b.y || b.z < a.y) return true;
    return D.thr > 0.0 && (double)min(a.x, b.x) / (double)max(a.x, b.x) < D.thr;
}
// The test for the rows i = perm[pa], j = perm[pb], read by position where
// the call's order has position records.
__device__ inline bool span_disjoint_at(const DevCorpus& D, uint32_t pa, uint32_t pb, uint32_t i,
                                        uint32_t j) {
    return D.pspan ? span_disjoint(D, D.pspan[pa], D.pspan[pb]) : span_disjoint(D, i, j);
}

// Coarse zone maps (bench/bench_tile.cpp:180-190, the per-pair filter in front
// of the cell): rows i and j share no bucket, so no position. Called by a
// whole wave; every lane gets the answer.
__device__ inline bool zone_disjoint(const DevCorpus& D, uint32_t i, uint32_t j, uint32_t lane) {
    const uint64_t* zi = D.zone + (uint64_t)i * D.zone_ld;
    const uint64_t* zj = D.zone + (uint64_t)j * D.zone_ld;
    uint64_t any = 0;
    for (uint32_t w = lane; w < D.zone_ld; w += 64) any |= zi[w] & zj[w];
    return __ballot(any != 0) == 0;
}

// The per-pair zone test pays where the tile pair's tiles share under half of
// the buckets (the gate); elsewhere nearly every pair shares one. Called by a
// whole wave.
__device__ inline bool zone_gate(const DevCorpus& D, uint32_t bi, uint32_t bj, uint32_t lane) {
    if (!D.zone || !D.tzone) return false;
    const uint64_t* a = D.tzone + (uint64_t)bi * D.zone_ld;
    const uint64_t* b = D.tzone + (uint64_t)bj * D.zone_ld;
    uint32_t c = 0;
    for (uint32_t w = lane; w < D.zone_ld; w += 64) c += __popcll(a[w] & b[w]);
    return 2u * wave_sum_u32(c) < D.zone_ld * 64u;
}

// ---- S x S at tile granularity: a hash join in LDS -------------------------
//
// For a block of up to 64 x 64 pairs of list rows, every pair's count is the
// number of positions the two rows share. Instead of intersecting 4,096 pairs
// separately, the I rows' positions go into an LDS hash table keyed by
// position whose value is a 64-bit mask of the I rows holding it; each J row's
// positions are then looked up once, and a hit adds 1 to count[i][j] for every
// i in the mask. Work is |I elements| + |J elements| + matches instead of
// sum over pairs. Positions are processed in universe chunks small enough that
// a chunk's I elements fill at most half the table.
__device__ inline uint32_t bit_len(uint32_t x) { return 32 - __clz(x); }

constexpr int kJoinBits = 12, kJoinSize = 1 << kJoinBits;
constexpr uint32_t kJoinEmpty = 0xFFFFFFFFu;

struct JoinSmem {
    uint32_t key[kJoinSize];
    uint32_t mlo[kJoinSize], mhi[kJoinSize];
    uint32_t cnt[64][65];
    uint32_t slot[kJoinSize / 2];   // table slot of each inserted element (for clearing)
    uint32_t start[2][64];     // first element of each row inside the chunk
    uint32_t pre[2][65];       // exclusive prefix of the rows' chunk counts
    uint32_t total_i;
};

__device__ inline uint32_t join_hash(uint32_t x) { return (x * 2654435761u) >> (32 - kJoinBits); }

// Row index of flattened element f: largest r with pre[r] <= f.
__device__ inline uint32_t join_row(const uint32_t* pre, uint32_t nrows, uint32_t f) {
    uint32_t lo = 0, hi = nrows;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= f) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ void ss_tile_join(const DevCorpus& D, JoinSmem& S, uint32_t ia, uint32_t ni,
                             uint32_t ja, uint32_t nj, bool diag,
                             uint64_t& local, uint64_t& pairs, uint64_t& skipped) {
    const uint32_t tid = threadIdx.x;
    for (uint32_t k = tid; k < 64 * 65; k += blockDim.x) (&S.cnt[0][0])[k] = 0;
    for (uint32_t k = tid; k < kJoinSize; k += blockDim.x) {
        S.key[k] = kJoinEmpty; S.mlo[k] = 0; S.mhi[k] = 0;
    }
    if (tid == 0) S.total_i = 0;
    __syncthreads();
    if (tid < ni) atomicAdd(&S.total_i, (uint32_t)(D.list_off[D.perm[ia + tid] + 1] - D.list_off[D.perm[ia + tid]]));
    __syncthreads();
    const uint64_t U = D.universe;
    const uint64_t target = kJoinSize / 4;
    uint64_t width = S.total_i <= target ? U : max<uint64_t>(1, U * target / S.total_i);
    uint64_t lo = 0;
    // Threads 0..63 follow the I rows, 64..127 the J rows: st = the row's first
    // element at or past lo (the previous chunk's end), en = past the chunk.
    uint32_t st = 0, en = 0;
    while (lo < U) {
        const uint64_t hi = min(U, lo + width);
        // Chunk spans of every row.
        if (tid < 128) {
            const uint32_t side = tid >> 6, r = tid & 63;
            const uint32_t nr = side ? nj : ni;
            uint32_t c = 0;
            if (r < nr) {
                const uint32_t row = D.perm[(side ? ja : ia) + r];
                const Span32 L = list_of(D, row);
                en = hi >= U ? L.n : st + gallop_lower_bound(L.p + st, L.n - st, (uint32_t)hi);
                c = en - st;
            }
            S.start[side][r] = st;
            S.pre[side][r + 1] = c;   // counts first; prefix below
        }
        __syncthreads();
        if (tid < 2) {
            uint32_t acc = 0;
            S.pre[tid][0] = 0;
            for (int r = 1; r <= 64; ++r) { acc += S.pre[tid][r]; S.pre[tid][r] = acc; }
        }
        __syncthreads();
        const uint32_t ei = S.pre[0][ni], ej = S.pre[1][nj];
        if (ei > kJoinSize / 2 && hi - lo > 1) {   // too full: halve the chunk
            width = (hi - lo) / 2;
            __syncthreads();
            continue;
        }
        for (uint32_t f = tid; f < ei; f += blockDim.x) {
            const uint32_t r = join_row(S.pre[0], ni, f);
            const Span32 L = list_of(D, D.perm[ia + r]);
            const uint32_t x = L.p[S.start[0][r] + (f - S.pre[0][r])];
            uint32_t h = join_hash(x);
            for (;;) {
                const uint32_t old = atomicCAS(&S.key[h], kJoinEmpty, x);
                if (old == kJoinEmpty || old == x) break;
                h = (h + 1) & (kJoinSize - 1);
            }
            if (r < 32) atomicOr(&S.mlo[h], 1u << r);
            else        atomicOr(&S.mhi[h], 1u << (r - 32));
            S.slot[f] = h;
        }
        __syncthreads();
        for (uint32_t f = tid; f < ej; f += blockDim.x) {
            const uint32_t r = join_row(S.pre[1], nj, f);
            const Span32 L = list_of(D, D.perm[ja + r]);
            const uint32_t x = L.p[S.start[1][r] + (f - S.pre[1][r])];
            uint32_t h = join_hash(x);
            for (;;) {
                const uint32_t k = S.key[h];
                if (k == kJoinEmpty) break;
                if (k == x) {
                    uint32_t m = S.mlo[h];
                    while (m) { const int b = __ffs(m) - 1; m &= m - 1; atomicAdd(&S.cnt[b][r], 1u); }
                    m = S.mhi[h];
                    while (m) { const int b = __ffs(m) - 1; m &= m - 1; atomicAdd(&S.cnt[32 + b][r], 1u); }
                    break;
                }
                h = (h + 1) & (kJoinSize - 1);
            }
        }
        __syncthreads();
        // Return the used slots to empty for the next chunk.
        for (uint32_t f = tid; f < ei; f += blockDim.x) {
            const uint32_t h = S.slot[f];
            S.key[h] = kJoinEmpty; S.mlo[h] = 0; S.mhi[h] = 0;
        }
        __syncthreads();
        lo = hi;
        st = en;
        if (ei < kJoinSize / 8) width = min<uint64_t>(U, width * 2);
    }
    for (uint32_t p = tid; p < ni * nj; p += blockDim.x) {
        const uint32_t a = p / nj, b = p % nj;
        if (diag && b <= a) continue;
        const uint32_t i = D.perm[ia + a], j = D.perm[ja + b];
        const uint32_t v = S.cnt[a][b];
        {
            // Counted as values, not as a choice between two counters (which
            // the compiler turns into a pointer into scratch memory).
            const uint64_t dis = span_disjoint_at(D, ia + a, ja + b, i, j) ? 1u : 0u;
            skipped += dis;
            pairs += 1u - dis;
        }
        local += v;
        emit_at(D, ia + a, ja + b, i, j, v);
    }
    __syncthreads();
}

// ---- B x S through bit-sliced bitmaps --------------------------------------
//
// A 64-row group of bitmap rows has a slice: per position, the mask of the
// group's rows that contain it; the eight groups of a slab (512 rows) keep
// their masks of a position side by side (64 bytes). A B x S item is a
// stripe: up to 64 rows of the probed side (lists) against up to 256 rows of
// the other side (one tile, 32 bytes per position) or 512 (a wide stripe over
// two tiles, the whole 64 bytes), whose groups all have slices. A wave takes
// one probed row and streams its list once; its lanes load 64 elements'
// masks and add 1 to the wave's LDS counter of every slab row set in them.
// One load per list element serves up to 512 pairs (the per-pair cell probes
// the bitmap once per element per pair); the counters are private to the
// wave, so no barrier. The loads are random and small: their number, not
// their bytes, bounds the stripes (a 64-byte load costs about 1.15x a 32-byte
// one), so the wide stripe halves the loads per pair.

// Orders one wave's LDS accesses across its lanes (no cross-wave effect).
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Counters per wave: one per slab row of the widest stripe.
constexpr uint32_t kStripeRows = 64 * kSlabGroups;

// Probe rows at sorted positions [p0, p0 + np), slab rows [s0, s0 + ns) (s0 a
// multiple of 64, G = ceil(ns / 64) groups, all with slices). Pairs are (A
// position, B position) with the probe rows on side A when probe_a; on a
// diagonal tile (same_tile) only pairs whose B position is the larger count.
template <int G>
__device__ void bs_stripe(const DevCorpus& D, uint32_t* __restrict__ wave_cnt, uint32_t p0,
                          uint32_t np, uint32_t s0, uint32_t ns, bool probe_a, bool same_tile,
                          uint64_t& local, uint64_t& pairs, uint64_t& skipped) {
    // 64-element batches a wave loads before counting (same bytes in flight
    // whatever the width).
    constexpr int F = G > 4 ? 2 : 4;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const uint32_t g0 = s0 >> 6;
    const uint64_t* base[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const uint32_t sl = D.slice_of[g0 + g];
        base[g] = D.slice + (uint64_t)(sl / kSlabGroups) * D.ld * 64 * kSlabGroups +
                  sl % kSlabGroups;
    }
    // The masks of a position in one or two 16-byte pairs: four groups of a
    // tile (32 bytes) or the eight of a slab (64 bytes), all in one slab.
    constexpr int V = G == 8 ? 4 : (G == 4 ? 2 : 0);
    const bool packed = V > 0 && (g0 % G) == 0;
    // This wave's counters, one per slab row: each lane adds 1 for every set
    // bit of its elements' masks (LDS atomics, a few per element at most
    // densities), then the row's counts are read back and cleared.
    uint32_t* cnt = wave_cnt + wave * kStripeRows;
    for (uint32_t q = lane; q < 64 * G; q += 64) cnt[q] = 0;
    // Lanes hand counters to each other through LDS (every lane adds to any
    // counter, lane k reads back counters k, k+64, ...): a wave-scope fence
    // and barrier order the clears, the adds and the read-back.
    wave_lds_sync();
    for (uint32_t r = wave; r < np; r += nwaves) {
        const Span32 L = list_of(D, D.perm[p0 + r]);
        for (uint32_t k0 = 0; k0 < L.n; k0 += 64 * F) {
            uint64_t m[F][G];
#pragma unroll
            for (int u = 0; u < F; ++u) {
                const uint32_t k = k0 + u * 64 + lane;
#pragma unroll
                for (int g = 0; g < G; ++g) m[u][g] = 0;
                if (k < L.n) {
                    const uint64_t x = (uint64_t)L.p[k] * kSlabGroups;
                    if (packed) {
                        const ulonglong2* q = reinterpret_cast<const ulonglong2*>(base[0] + x);
#pragma unroll
                        for (int v = 0; v < V; ++v) {
                            const ulonglong2 w = q[v];
                            m[u][(2 * v) % G] = w.x;
                            m[u][(2 * v + 1) % G] = w.y;
                        }
                    } else {
#pragma unroll
                        for (int g = 0; g < G; ++g) m[u][g] = base[g][x];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < F; ++u)
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    unsigned long long x = m[u][g];
                    while (x) {
                        const uint32_t a = (uint32_t)__ffsll(x) - 1;
                        x &= x - 1;
                        atomicAdd(&cnt[64 * g + a], 1u);
                    }
                }
        }
        wave_lds_sync();
        // The probe row's span record (one for the wave) and each slab row's,
        // read by position: 16 contiguous bytes per lane.
        const uint4 sp_r = D.pspan ? D.pspan[p0 + r] : uint4{0, 0, 0, 0};
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const uint32_t sa = 64 * g + lane;
            const uint32_t v = cnt[sa];
            cnt[sa] = 0;
            if (sa >= ns) continue;
            const uint32_t pa = probe_a ? p0 + r : s0 + sa, pb = probe_a ? s0 + sa : p0 + r;
            if (same_tile && pb <= pa) continue;
            const uint32_t i = D.perm[pa], j = D.perm[pb];
            {
                const bool d = D.pspan ? span_disjoint(D, sp_r, D.pspan[s0 + sa])
                                       : span_disjoint(D, i, j);
                const uint64_t dis = d ? 1u : 0u;
                skipped += dis;
                pairs += 1u - dis;
            }
            local += v;
            emit_at(D, pa, pb, i, j, v);
        }
        wave_lds_sync();
    }
}

// ---- R x R at tile granularity: an interval join in LDS --------------------
//
// For a block of up to 64 x 64 pairs of run rows, count[i][j] = sum over run
// pairs of their overlap. The universe is walked in chunks of kRunBuckets
// buckets of 2^sh positions. Each "I" run clipped to the chunk sets its row's
// bit in the 64-bit cover mask of every bucket it covers whole and leaves an
// entry (row, first, last offset) in the list of a bucket it covers in part.
// Each "J" run then visits the buckets it touches: every row in a bucket's
// mask gains the J run's overlap with the bucket, every entry its overlap with
// the entry. Work is (I runs + J runs) x buckets touched + matches instead of
// a search per run per pair. The chunk width adapts so that a chunk's I runs
// fill about three quarters of the entry pool; a chunk whose entries overflow the pool is
// redone at half the width (at width 1 every cover is whole: no entries).
constexpr uint32_t kRunBuckets = 1024, kRunEntries = 2048, kRunNoEntry = 0xFFFFu;
constexpr uint32_t kRunMaxShift = 16;   // entry offsets fit 16 bits
constexpr int kRunFetch = 4;            // runs each thread loads before using them
constexpr uint32_t kRunLong = 256;      // multi-bucket J runs queued per chunk (more: visited in place)

struct RunJoinSmem {
    uint32_t cnt[64][65];
    unsigned long long cover[kRunBuckets];
    uint32_t head[kRunBuckets];          // first entry of each bucket's list
    uint16_t next[kRunEntries];
    uint32_t span[kRunEntries];          // first | last << 16, offsets in the bucket
    uint8_t row[kRunEntries];
    uint64_t first[2][64];               // run-array index of each row's first run in the chunk
    uint32_t pre[2][65];             // exclusive prefix of the rows' chunk run counts
    uint3 long_run[kRunLong];            // queued J runs over several buckets: row, start, end
    uint32_t n_entries, n_long;
};

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ inline uint32_t wave_scan_u32(uint32_t v) {
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += t;
    }
    return v;
}

// swap: the I side (inserted) is the block's J rows; counts are stored by
// (I row, J row) and read back transposed.
__device__ void rr_tile_join(const DevCorpus& D, RunJoinSmem& S, uint32_t ia, uint32_t ni,
                             uint32_t ja, uint32_t nj, bool diag, bool swap,
                             uint64_t& local, uint64_t& pairs, uint64_t& skipped) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t xs0 = swap ? ja : ia, nx = swap ? nj : ni;   // inserted side
    const uint32_t ys0 = swap ? ia : ja, ny = swap ? ni : nj;   // queried side
    for (uint32_t k = tid; k < 64 * 65; k += blockDim.x) (&S.cnt[0][0])[k] = 0;
    for (uint32_t k = tid; k < kRunBuckets; k += blockDim.x) { S.cover[k] = 0; S.head[k] = kRunNoEntry; }
    if (tid == 0) { S.n_entries = 0; S.n_long = 0; }
    // Starting width: about 3/4 kRunEntries I runs per chunk at uniform spread.
    uint64_t total = 0;
    if (wave == 0) {
        const uint32_t c = lane < nx ? runs_of(D, D.perm[xs0 + lane]).n : 0u;
        total = wave_sum_u64(c);
    }
    __shared__ uint64_t s_total;
    if (tid == 0) s_total = total;
    __syncthreads();
    const uint64_t U = D.universe;
    uint32_t sh = 0;
    {
        const uint64_t want = s_total ? U * (kRunEntries * 3 / 4) / s_total : U;
        while (sh < kRunMaxShift && ((uint64_t)kRunBuckets << (sh + 1)) <= want) ++sh;
    }
    // Threads 0..63 follow the I rows, 64..127 the J rows: st = the row's first
    // run ending after lo (runs are disjoint and sorted by start, so their ends
    // are sorted too), en = one past its last run starting before hi.
    const uint32_t side = (tid >> 6) & 1, r_own = tid & 63;
    const bool follows = tid < 128 && r_own < (side ? ny : nx);
    RunsV own{nullptr, nullptr, 0};
    uint64_t own_off = 0;
    if (follows) {
        const uint32_t row = D.perm[(side ? ys0 : xs0) + r_own];
        own = runs_of(D, row);
        own_off = D.run_off[row];
    }
    uint32_t st = 0, en = 0;
    uint64_t lo = 0;
    while (lo < U) {
        const uint64_t hi = min(U, lo + ((uint64_t)kRunBuckets << sh));
        if (tid < 128) {
            uint32_t c = 0;
            if (follows) {
                en = hi >= U ? own.n : st + gallop_lower_bound(own.s + st, own.n - st, (uint32_t)hi);
                S.first[side][r_own] = own_off + st;
                c = en - st;
            }
            const uint32_t inc = wave_scan_u32(c);
            S.pre[side][r_own + 1] = inc;
            if (r_own == 0) S.pre[side][0] = 0;
        }
        __syncthreads();
        const uint32_t ex = S.pre[0][nx], ey = S.pre[1][ny];
        const uint32_t w = 1u << sh;
        // Each thread fetches kRunFetch runs (clipped to the chunk, relative
        // to lo) before working on any, so their loads overlap.
        auto fetch = [&](int sd, uint32_t nr, uint32_t f, uint32_t& r, uint32_t& rs, uint32_t& re) {
            r = join_row(S.pre[sd], nr, f);
            const uint64_t g = S.first[sd][r] + (f - S.pre[sd][r]);
            rs = (uint32_t)(max<uint64_t>(D.run_start[g], lo) - lo);
            re = (uint32_t)(min<uint64_t>(D.run_end[g], hi) - lo);
        };
        // Insert the I runs.
        for (uint32_t f0 = tid; f0 < ex; f0 += kRunFetch * blockDim.x) {
            uint32_t rr[kRunFetch], rss[kRunFetch], rse[kRunFetch];
#pragma unroll
            for (int u = 0; u < kRunFetch; ++u)
                if (f0 + u * blockDim.x < ex) fetch(0, nx, f0 + u * blockDim.x, rr[u], rss[u], rse[u]);
#pragma unroll
            for (int u = 0; u < kRunFetch; ++u) {
                if (f0 + u * blockDim.x >= ex) break;
                const uint32_t r = rr[u], rs = rss[u], re = rse[u];
                const uint32_t kb = rs >> sh, ke = (re - 1) >> sh;
                const unsigned long long bit = 1ull << r;
                auto part = [&](uint32_t b, uint32_t first, uint32_t last) {
                    const uint32_t x = atomicAdd(&S.n_entries, 1u);
                    if (x >= kRunEntries) return;
                    S.span[x] = first | (last << 16);
                    S.row[x] = (uint8_t)r;
                    S.next[x] = (uint16_t)atomicExch(&S.head[b], x);
                };
                if (kb == ke) {
                    if (re - rs == w) atomicOr(&S.cover[kb], bit);
                    else part(kb, rs & (w - 1), (re - 1) & (w - 1));
                } else {
                    if (rs & (w - 1)) part(kb, rs & (w - 1), w - 1);
                    else atomicOr(&S.cover[kb], bit);
                    for (uint32_t b = kb + 1; b < ke; ++b) atomicOr(&S.cover[b], bit);
                    if (re & (w - 1)) part(ke, 0, (re - 1) & (w - 1));
                    else atomicOr(&S.cover[ke], bit);
                }
            }
        }
        __syncthreads();
        const bool overflow = S.n_entries > kRunEntries;
        if (!overflow) {
            // Query the J runs. A run's work in one bucket: every row in the
            // bucket's cover mask gains the run's overlap with the bucket,
            // every entry its overlap with the entry.
            auto visit = [&](uint32_t r, uint32_t rs, uint32_t re, uint32_t b) {
                const uint32_t b0 = b << sh;
                const uint32_t qs = max(rs, b0) - b0, qe = min(re, b0 + w) - b0;   // [qs, qe) in the bucket
                unsigned long long m = S.cover[b];
                while (m) {
                    const int a = __ffsll(m) - 1;
                    m &= m - 1;
                    atomicAdd(&S.cnt[a][r], qe - qs);
                }
                for (uint32_t x = S.head[b]; x != kRunNoEntry; x = S.next[x]) {
                    const uint32_t es = S.span[x] & 0xFFFFu, el = S.span[x] >> 16;
                    const uint32_t o0 = max(qs, es), o1 = min(qe, el + 1);
                    if (o0 < o1) atomicAdd(&S.cnt[S.row[x]][r], o1 - o0);
                }
            };
            // Runs within one bucket are visited here; a run over several
            // buckets is queued (a lane walking its buckets would hold back
            // its whole wave) and visited below by a wave, one bucket a lane.
            for (uint32_t f0 = tid; f0 < ey; f0 += kRunFetch * blockDim.x) {
                uint32_t rr[kRunFetch], rss[kRunFetch], rse[kRunFetch];
#pragma unroll
                for (int u = 0; u < kRunFetch; ++u)
                    if (f0 + u * blockDim.x < ey) fetch(1, ny, f0 + u * blockDim.x, rr[u], rss[u], rse[u]);
#pragma unroll
                for (int u = 0; u < kRunFetch; ++u) {
                    if (f0 + u * blockDim.x >= ey) break;
                    const uint32_t r = rr[u], rs = rss[u], re = rse[u];
                    const uint32_t kb = rs >> sh, ke = (re - 1) >> sh;
                    if (kb != ke) {
                        const uint32_t q = atomicAdd(&S.n_long, 1u);
                        if (q < kRunLong) {
                            S.long_run[q] = make_uint3(r, rs, re);
                            continue;
                        }
                    }
                    for (uint32_t b = kb; b <= ke; ++b) visit(r, rs, re, b);
                }
            }
            __syncthreads();
            const uint32_t n_long = min(S.n_long, kRunLong);
            for (uint32_t q = wave; q < n_long; q += blockDim.x >> 6) {
                const uint3 L = S.long_run[q];
                const uint32_t kb = L.y >> sh, ke = (L.z - 1) >> sh;
                for (uint32_t b = kb + lane; b <= ke; b += 64) visit(L.x, L.y, L.z, b);
            }
        }
        __syncthreads();
        for (uint32_t k = tid; k < kRunBuckets; k += blockDim.x) { S.cover[k] = 0; S.head[k] = kRunNoEntry; }
        __syncthreads();
        if (overflow) {
            // Redo the chunk at half the width (sh > 0 here: width-1 buckets
            // take no entries).
            if (tid == 0) S.n_entries = 0;
            --sh;
            __syncthreads();
            continue;
        }
        const uint32_t used = S.n_entries;
        __syncthreads();
        if (tid == 0) { S.n_entries = 0; S.n_long = 0; }
        // The last run starting before hi may continue past it.
        if (follows) st = (en > st && (uint64_t)own.e[en - 1] > hi) ? en - 1 : en;
        lo = hi;
        if (used < kRunEntries / 8 && sh < kRunMaxShift) ++sh;
        else if (used > kRunEntries * 15 / 16 && sh > 0) --sh;
    }
    for (uint32_t p = tid; p < ni * nj; p += blockDim.x) {
        const uint32_t a = p / nj, b = p % nj;
        if (diag && b <= a) continue;
        const uint32_t i = D.perm[ia + a], j = D.perm[ja + b];
        const uint32_t v = swap ? S.cnt[b][a] : S.cnt[a][b];
        {
            // Counted as values, not as a choice between two counters (which
            // the compiler turns into a pointer into scratch memory).
            const uint64_t dis = span_disjoint_at(D, ia + a, ja + b, i, j) ? 1u : 0u;
            skipped += dis;
            pairs += 1u - dis;
        }
        local += v;
        emit_at(D, ia + a, ja + b, i, j, v);
    }
    __syncthreads();
}

// ---- R x R and S x S with one row staged in LDS ------------------------------
//
// Per-pair search in global memory is a chain of dependent loads (a binary
// search per run or element). Here the workgroup stages one row of the tile's
// "staged" side in LDS (its run starts and ends, or its sorted positions) and
// its waves take the other side's rows one by one: every lane streams that
// row's runs or positions from global memory (coalesced) and searches the
// LDS copy. The staged side is the one with more runs (elements) per row, so
// the streamed side is the shorter one; a staged row must fit kStageItems.
constexpr uint32_t kStageItems = 4096;

// Bucket index over the staged row: the universe in nb buckets of 2^shift
// positions (nb the power of two at or above the row's item count, at most
// kStageBuckets); idx[k] = first staged item whose key (run end for R x R,
// position for S x S) is above k << shift, so a lookup bisects one bucket's
// items instead of the whole row, and the index costs O(items) to build.
constexpr uint32_t kStageBuckets = 4096;

struct StageSmem {
    uint32_t a[kStageItems];   // run starts / sorted positions
    uint32_t b[kStageItems];   // run ends (R x R)
    uint16_t idx[kStageBuckets];   // item indices <= kStageItems fit 16 bits
};

// Fills S.idx for the staged keys key[0..n) (ascending; run ends, or
// positions + 1 so that "above" means "at or above the position").
__device__ inline void stage_index(StageSmem& S, const uint32_t* key, uint32_t n, uint32_t nb,
                                   uint32_t shift, bool plus_one) {
    for (uint32_t p = threadIdx.x; p <= n; p += blockDim.x) {
        // Buckets whose first answer is p: prev_key <= k << shift < key[p].
        const uint64_t prev = p ? (uint64_t)key[p - 1] + plus_one : 0;
        const uint64_t lo = p ? (prev + (1ull << shift) - 1) >> shift : 0;
        const uint64_t hi = p < n ? (((uint64_t)key[p] + plus_one - 1) >> shift) + 1 : (uint64_t)nb;
        for (uint64_t k = lo; k < hi && k < nb; ++k) S.idx[k] = (uint16_t)p;
    }
}

template <int CELL>
__device__ void tile_staged(const DevCorpus& D, StageSmem& S, uint32_t ia, uint32_t ni,
                            uint32_t ja, uint32_t nj, bool diag, bool stage_j,
                            uint64_t& local, uint64_t& pairs, uint64_t& skipped) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
    const uint32_t xs0 = stage_j ? ja : ia, nx = stage_j ? nj : ni;
    const uint32_t ys0 = stage_j ? ia : ja, ny = stage_j ? ni : nj;
    for (uint32_t u = 0; u < nx; ++u) {
        const uint32_t x = D.perm[xs0 + u];
        __syncthreads();   // the previous staged row is no longer read
        uint32_t nx_items;
        if constexpr (CELL == STORM_HIP_CELL_RR) nx_items = runs_of(D, x).n;
        else nx_items = list_of(D, x).n;
        // nb buckets of 2^shift positions cover the universe.
        uint32_t nb = 1, shift = 0;
        while (nb < nx_items && nb < kStageBuckets) nb <<= 1;
        while (((uint64_t)nb << shift) < D.universe) ++shift;
        if constexpr (CELL == STORM_HIP_CELL_RR) {
            const RunsV R = runs_of(D, x);
            for (uint32_t k = tid; k < R.n; k += blockDim.x) { S.a[k] = R.s[k]; S.b[k] = R.e[k]; }
            stage_index(S, R.e, R.n, nb, shift, false);
        } else {
            const Span32 L = list_of(D, x);
            for (uint32_t k = tid; k < L.n; k += blockDim.x) S.a[k] = L.p[k];
            stage_index(S, L.p, L.n, nb, shift, true);
        }
        __syncthreads();
        // First staged item at or after the bucket of `key` whose value is
        // above `key` (run ends for R x R, positions + 1 for S x S): a
        // bisection inside the key's bucket.
        auto lookup = [&](const uint32_t* vals, uint32_t key, uint32_t plus_one) {
            const uint32_t k = key >> shift;
            uint32_t lo = S.idx[k];
            uint32_t hi = k + 1 < nb ? S.idx[k + 1] : nx_items;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (vals[mid] + plus_one <= key) lo = mid + 1; else hi = mid;
            }
            return lo;
        };
        for (uint32_t v = wave; v < ny; v += nwaves) {
            // Tile coordinates of the pair: (row in I, row in J).
            const uint32_t ti = stage_j ? v : u, tj = stage_j ? u : v;
            if (diag && tj <= ti) continue;
            const uint32_t y = D.perm[ys0 + v];
            uint32_t cnt = 0;
            const bool dis = span_disjoint(D, x, y);
            if (lane == 0) { skipped += dis ? 1u : 0u; pairs += dis ? 0u : 1u; }
            if (!dis) {
                // Lanes take the streamed row's items 64 apart (coalesced
                // loads) and look each up in the staged row's bucket index.
                uint32_t c = 0;
                if constexpr (CELL == STORM_HIP_CELL_RR) {
                    const RunsV Q = runs_of(D, y);
#pragma unroll 4
                    for (uint32_t k = lane; k < Q.n; k += 64) {
                        const uint32_t rs = Q.s[k], re = Q.e[k];
                        // first staged run ending after rs
                        const uint32_t q = lookup(S.b, rs, 0u);
                        for (uint32_t p = q; p < nx_items && S.a[p] < re; ++p)
                            c += min(re, S.b[p]) - max(rs, S.a[p]);
                    }
                } else {
                    const Span32 Q = list_of(D, y);
#pragma unroll 4
                    for (uint32_t k = lane; k < Q.n; k += 64) {
                        const uint32_t e = Q.p[k];
                        const uint32_t q = lookup(S.a, e, 1u);   // first staged position >= e
                        c += (q < nx_items && S.a[q] == e) ? 1u : 0u;
                    }
                }
                cnt = wave_sum_u32(c);
            }
            if (lane == 0) {
                local += cnt;
                emit_pair(D, x, y, cnt);
            }
        }
    }
    __syncthreads();
}

// Which side of a tile pair to stage in LDS (-1: neither fits): the side with
// more items per row, if its largest row fits.
template <int CELL>
__device__ inline int stage_side(const TileMeta& A, const TileMeta& B) {
    const uint32_t ma = CELL == STORM_HIP_CELL_RR ? A.max_runs : A.max_card;
    const uint32_t mb = CELL == STORM_HIP_CELL_RR ? B.max_runs : B.max_card;
    const uint32_t ua = CELL == STORM_HIP_CELL_RR ? A.runs : A.card;
    const uint32_t ub = CELL == STORM_HIP_CELL_RR ? B.runs : B.card;
    const bool fa = ma <= kStageItems, fb = mb <= kStageItems;
    if (fa && fb) return ub > ua ? 1 : 0;
    return fa ? 0 : fb ? 1 : -1;
}

// The join wins over per-pair search unless the J rows dwarf the I rows.
__device__ inline bool ss_use_join(const TileMeta& A, const TileMeta& B) {
    const double ci = A.card, cj = B.card;
    const double rows = (double)A.rows * B.rows;
    const double per_pair = rows * fmin(ci, cj) * (double)bit_len(max(A.card, B.card) + 2);
    const double join = 3.0 * (A.rows * ci + B.rows * cj) + A.rows * ci / (kJoinSize / 4.0) * 2200.0;
    return join < per_pair;
}

// Interval join against staged per-pair search for R x R: the join touches
// every run of both sides once (plus per-chunk row bookkeeping), the search
// every run of the shorter row once per pair, at a bisection each.
__device__ inline bool rr_use_join(const TileMeta& A, const TileMeta& B) {
    // Per 64 x 64 block, the unit both run in.
    const double ra = A.runs, rb = B.runs, na = min(A.rows, 64u), nb = min(B.rows, 64u);
    const double per_pair = na * nb * fmin(ra, rb) * (double)(bit_len(max(A.runs, B.runs) + 2) + 2);
    const double join = 8.0 * (na * ra + nb * rb) + (fmin(na * ra, nb * rb) / (kRunEntries / 2.0) + 1.0) * 2000.0;
    return join < per_pair;
}

}  // namespace

template <int CELL>
__global__ __launch_bounds__(256) void cell_kernel(const DevCorpus* __restrict__ Ds, DevModel M,
                                                   const QItem* __restrict__ queue,
                                                   DevCounters* __restrict__ ctr,
                                                   unsigned long long* __restrict__ tile_cycles) {
    __shared__ uint32_t s_t;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t pairs = 0, skipped = 0;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_t = atomicAdd(&ctr->qhead[CELL], 1u);
        __syncthreads();
        const uint32_t t = s_t;
        if (t >= ctr->qcount[CELL]) break;
        const uint64_t t0 = wall_clock64();
        const QItem e = queue[t];
        const DevCorpus D = Ds[e.inst];
        const uint32_t n = D.n, tile = D.tile;
        uint64_t local = 0;
        // A shadow item (probe runner-up) leaves the call's pair counts alone.
        const uint64_t pairs0 = pairs, skipped0 = skipped;
        uint32_t i0 = e.bi * tile, j0 = e.bj * tile;
        uint32_t ni = min(n, i0 + tile) - i0, nj = min(n, j0 + tile) - j0;
        const bool wide = e.sub && (e.sub >> 24) == kWideSide;
        if (wide) {
            // Both bitmap tiles bj, bj + 1 (full, so 512 rows from j0).
            nj = 2 * tile;
        }
        if (e.sub) {
            // One block of the tile (empty when it lies past the last row), or
            // a stripe: one side whole (or two tiles whole: wide).
            const uint32_t sb = e.sub & 0xFFFFu;
            const uint32_t xi = (e.sub >> 16) & 0xFFu, xj = e.sub >> 24;
            if (xi != kWholeSide) {
                const uint32_t oi = xi * sb;
                ni = oi < ni ? min(ni - oi, sb) : 0u;
                i0 += oi;
            }
            if (xj != kWholeSide && xj != kWideSide) {
                const uint32_t oj = xj * sb;
                nj = oj < nj ? min(nj - oj, sb) : 0u;
                j0 += oj;
            }
        }
        // Pairs i < j only where the block straddles the diagonal (a stripe of
        // a diagonal tile keeps the pairs whose column position is larger).
        const bool diag = i0 == j0;
        const bool same_tile = e.bi == e.bj;
        const uint32_t np = ni * nj;
        // Lane mode when the sparser tile's rows are tiny.
        const TileMeta& A = D.tm[e.bi];
        const bool lane_mode = CELL == STORM_HIP_CELL_EMPTY || (A.max_card <= 48 && A.max_runs <= 48);
        bool joined = false;
        if constexpr (CELL == STORM_HIP_CELL_BS) {
            const TileMeta& B = D.tm[e.bj];
            // A stripe names its probed side; a whole tile or block probes the
            // side bs_probe_a picks.
            bool probe_a = bs_probe_a(A, B);
            if (e.sub && (e.sub >> 24) >= kWideSide) probe_a = true;
            else if (e.sub && ((e.sub >> 16) & 0xFFu) == kWholeSide) probe_a = false;
            const uint32_t s0 = probe_a ? j0 : i0, ns = probe_a ? nj : ni;
            const uint32_t p0 = probe_a ? i0 : j0, np_ = probe_a ? ni : nj;
            // A wide stripe covers 512 rows, every other stripe one tile.
            bool ok = D.slice != nullptr && (s0 & 63) == 0 && ns > 0 && np_ > 0 &&
                      (wide ? ns == 512 : ns <= 256);
            for (uint32_t g = 0; ok && g < (ns + 63) / 64; ++g)
                ok = D.slice_of[(s0 >> 6) + g] != kNoBitmap;
            const uint32_t probe_card = probe_a ? A.card : B.card;
            // The select kernel queues wide stripes only where both halves
            // pass this test.
            if (ok && (wide || bs_stripes_pay(probe_card, A, B))) {
                joined = true;
                const bool same = e.bi == e.bj;
                __shared__ uint32_t wave_cnt[4 * kStripeRows];   // per wave (4 waves)
                switch ((ns + 63) / 64) {
                    case 1: bs_stripe<1>(D, wave_cnt, p0, np_, s0, ns, probe_a, same, local, pairs, skipped); break;
                    case 2: bs_stripe<2>(D, wave_cnt, p0, np_, s0, ns, probe_a, same, local, pairs, skipped); break;
                    case 3: bs_stripe<3>(D, wave_cnt, p0, np_, s0, ns, probe_a, same, local, pairs, skipped); break;
                    case 4: bs_stripe<4>(D, wave_cnt, p0, np_, s0, ns, probe_a, same, local, pairs, skipped); break;
                    default: bs_stripe<8>(D, wave_cnt, p0, np_, s0, ns, probe_a, same, local, pairs, skipped); break;
                }
            }
        }
        if constexpr (CELL == STORM_HIP_CELL_SS) {
            // The join and the staged search never run together: one LDS area.
            __shared__ union { JoinSmem js; StageSmem ss; } sm;
            if (ss_use_join(A, D.tm[e.bj])) {
                joined = true;
                for (uint32_t si = 0; si < ni; si += 64)
                    for (uint32_t sj = diag ? si : 0; sj < nj; sj += 64)
                        ss_tile_join(D, sm.js, i0 + si, min(64u, ni - si), j0 + sj,
                                     min(64u, nj - sj), diag && si == sj, local, pairs,
                                     skipped);
            } else if (!lane_mode) {
                const int side = stage_side<CELL>(A, D.tm[e.bj]);
                if (side >= 0) {
                    joined = true;
                    tile_staged<CELL>(D, sm.ss, i0, ni, j0, nj, diag, side == 1, local, pairs,
                                      skipped);
                }
            }
        }
        if constexpr (CELL == STORM_HIP_CELL_RR) {
            __shared__ union { RunJoinSmem rj; StageSmem ss; } sm;
            const TileMeta& B = D.tm[e.bj];
            if (rr_use_join(A, B)) {
                joined = true;
                // Insert the side with fewer runs: fewer chunks.
                const bool swap = !diag && (uint64_t)B.runs * B.rows < (uint64_t)A.runs * A.rows;
                for (uint32_t si = 0; si < ni; si += 64)
                    for (uint32_t sj = diag ? si : 0; sj < nj; sj += 64)
                        rr_tile_join(D, sm.rj, i0 + si, min(64u, ni - si), j0 + sj,
                                     min(64u, nj - sj), diag && si == sj, swap && !(diag && si == sj),
                                     local, pairs, skipped);
            } else if (!lane_mode) {
                const int side = stage_side<CELL>(A, B);
                if (side >= 0) {
                    joined = true;
                    tile_staged<CELL>(D, sm.ss, i0, ni, j0, nj, diag, side == 1, local, pairs,
                                      skipped);
                }
            }
        }
        if (joined) {
        } else if (lane_mode) {
            for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) {
                const uint32_t a = i0 + p / nj, b = j0 + p % nj;
                if (same_tile && b <= a) continue;
                const uint32_t i = D.perm[a], j = D.perm[b];
                uint32_t cnt = 0;
                const bool dis = CELL == STORM_HIP_CELL_EMPTY || span_disjoint(D, i, j);
                skipped += dis ? 1u : 0u;
                pairs += dis ? 0u : 1u;
                if (!dis) {
                    cnt = cell_partial<CELL>(D, M, i, j, 0, 1);
                    if (CELL == STORM_HIP_CELL_CB) cnt = cb_base(D, i, j) - cnt;
                }
                local += cnt;
                emit_pair(D, i, j, cnt);
            }
        } else {
            const bool gate = zone_gate(D, e.bi, e.bj, lane);
            for (uint32_t p = wave; p < np; p += blockDim.x / 64) {
                const uint32_t a = i0 + p / nj, b = j0 + p % nj;
                if (same_tile && b <= a) continue;
                const uint32_t i = D.perm[a], j = D.perm[b];
                uint32_t cnt = 0;
                const bool dis = span_disjoint(D, i, j) || (gate && zone_disjoint(D, i, j, lane));
                if (lane == 0) { skipped += dis ? 1u : 0u; pairs += dis ? 0u : 1u; }
                if (!dis) {
                    cnt = wave_sum_u32(cell_partial<CELL>(D, M, i, j, lane, 64));
                    if (CELL == STORM_HIP_CELL_CB) cnt = cb_base(D, i, j) - cnt;
                }
                if (lane == 0) {
                    local += cnt;
                    emit_pair(D, i, j, cnt);
                }
            }
        }
        if (D.shadow) { pairs = pairs0; skipped = skipped0; }
        local = wave_sum_u64(local);
        if (lane == 0 && local) atomicAdd(D.sum, (unsigned long long)local);
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint64_t now = wall_clock64();
            if (tile_cycles) tile_cycles[t] = now - t0;
            if (D.finish) atomicMax(D.finish, (unsigned long long)now);
        }
    }
    pairs = wave_sum_u64(pairs);
    skipped = wave_sum_u64(skipped);
    if (lane == 0) {
        if (pairs) atomicAdd(&ctr->cell_pairs[CELL], (unsigned long long)pairs);
        if (skipped) {
            atomicAdd(&ctr->cell_pairs[STORM_HIP_CELL_EMPTY], (unsigned long long)skipped);
            atomicAdd(&ctr->skipped, (unsigned long long)skipped);
        }
    }
}

// One row's metadata as a one-row tile, for per-pair pricing.
__device__ inline TileMeta row_tile(const DevCorpus& D, uint32_t r) {
    TileMeta m;
    m.card = meta(D, r, 0); m.runs = meta(D, r, 1); m.nw = meta(D, r, 2); m.nnz = meta(D, r, 3);
    m.first = m.card ? meta(D, r, 4) : 0; m.last = meta(D, r, 5);
    m.has_rank = meta(D, r, 6); m.has_occ = meta(D, r, 7);
    m.all_bitmap = D.bm_row[r] != kNoBitmap;
    m.all_comp = 2ull * m.card > D.universe;
    m.any_empty = m.card == 0;
    m.max_card = m.min_card = m.card; m.max_runs = m.runs;
    m.max_ewah = (uint32_t)(D.mk_off[r + 1] - D.mk_off[r]);
    m.rows = 1;
    m.spread = 1.0;
    return m;
}

// One pair's B x B with the zone-map plan (bb_occ_sel, kernels/cell_bb.cpp:
// 537-550): when both rows carry a zone map (one bit per 8-word bin), the
// bins live on both sides are counted first; none settles the pair at 0,
// fewer than a third of all bins (kOccPlanSelectivity, :522) visits only
// those bins, else every word. Lane `first` of `stride` (a whole wave:
// first = lane, stride = 64, when the plan is taken).
__device__ uint32_t bb_pair_partial(const DevCorpus& D, uint32_t i, uint32_t j, uint32_t first,
                                    uint32_t stride) {
    const uint64_t* a = bitmap_of(D, i);
    const uint64_t* b = bitmap_of(D, j);
    const uint64_t* oa = occ_of(D, i);
    const uint64_t* ob = occ_of(D, j);
    uint32_t c = 0;
    if (oa && ob && stride == 64) {
        const uint32_t bins = (D.nw + 7) / 8;
        uint32_t live = 0;
        for (uint32_t k = first; k < D.occ_ld; k += 64) live += __popcll(oa[k] & ob[k]);
        live = wave_sum_u32(live);
        if (live == 0) return 0;
        if ((double)live < 0.33 * (double)bins) {
            // Each lane takes the live bins of its own zone-map words.
            for (uint32_t k = first; k < D.occ_ld; k += 64) {
                uint64_t m = oa[k] & ob[k];
                while (m) {
                    const uint32_t bin = k * 64 + (uint32_t)__ffsll((unsigned long long)m) - 1;
                    m &= m - 1;
                    const uint32_t w0 = bin * 8, w1 = min(w0 + 8, D.nw);
                    for (uint32_t w = w0; w < w1; ++w) c += __popcll(a[w] & b[w]);
                }
            }
            return c;
        }
    }
    for (uint32_t k = first; k < D.nw; k += stride) c += __popcll(a[k] & b[k]);
    return c;
}

// Lane `first` of `stride` of one pair's count under a runtime-chosen cell
// (B x B: AND + popcount over the words; the rest as their kernels count).
__device__ uint32_t pair_partial(int cell, const DevCorpus& D, const DevModel& M, uint32_t i,
                                 uint32_t j, uint32_t first, uint32_t stride) {
    switch (cell) {
        case STORM_HIP_CELL_BB: return bb_pair_partial(D, i, j, first, stride);
        case STORM_HIP_CELL_BS: return cell_partial<STORM_HIP_CELL_BS>(D, M, i, j, first, stride);
        case STORM_HIP_CELL_BR: return cell_partial<STORM_HIP_CELL_BR>(D, M, i, j, first, stride);
        case STORM_HIP_CELL_BW: return cell_partial<STORM_HIP_CELL_BW>(D, M, i, j, first, stride);
        case STORM_HIP_CELL_SS: return cell_partial<STORM_HIP_CELL_SS>(D, M, i, j, first, stride);
        case STORM_HIP_CELL_SR: return cell_partial<STORM_HIP_CELL_SR>(D, M, i, j, first, stride);
        case STORM_HIP_CELL_SW: return cell_partial<STORM_HIP_CELL_SW>(D, M, i, j, first, stride);
        case STORM_HIP_CELL_RR: return cell_partial<STORM_HIP_CELL_RR>(D, M, i, j, first, stride);
        case STORM_HIP_CELL_RW: return cell_partial<STORM_HIP_CELL_RW>(D, M, i, j, first, stride);
        default: return cell_partial<STORM_HIP_CELL_WW>(D, M, i, j, first, stride);
    }
}

// Refined tiles (Policy::Refine, kernels/storm_allpairs.cpp:563-588 and
// :636-641): per pair, the cheaper of the tile's two candidates as priced on
// the pair's own rows (denser row first), one wavefront per pair, the span
// test first. Per-pair tiles (Policy::PerPair / Oracle, :645-660, candidates
// both kPerPairCell): select_pairing on the pair's rows over every cell the
// two rows can run. Items are whole tiles or kSplitRows blocks; the item's
// inst carries the candidates above kRefineShift.
__global__ __launch_bounds__(256) void refine_kernel(const DevCorpus* __restrict__ Ds, DevModel M,
                                                     const QItem* __restrict__ queue,
                                                     DevCounters* __restrict__ ctr,
                                                     unsigned long long* __restrict__ tile_cycles) {
    __shared__ uint32_t s_t;
    __shared__ unsigned long long s_pairs[kNumCells];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x < kNumCells) s_pairs[threadIdx.x] = 0;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_t = atomicAdd(&ctr->qhead[kQueueRefine], 1u);
        __syncthreads();
        const uint32_t t = s_t;
        if (t >= ctr->qcount[kQueueRefine]) break;
        const uint64_t t0 = wall_clock64();
        const QItem e = queue[t];
        const uint32_t inst = e.inst & ((1u << kRefineShift) - 1);
        const int first = (e.inst >> kRefineShift) & 15, second = (e.inst >> (kRefineShift + 4)) & 15;
        const DevCorpus D = Ds[inst];
        const uint32_t n = D.n, tile = D.tile;
        uint32_t i0 = e.bi * tile, j0 = e.bj * tile;
        uint32_t ni = min(n, i0 + tile) - i0, nj = min(n, j0 + tile) - j0;
        if (e.sub) {
            const uint32_t sb = e.sub & 0xFFFFu;
            const uint32_t oi = ((e.sub >> 16) & 0xFFu) * sb, oj = (e.sub >> 24) * sb;
            ni = oi < ni ? min(ni - oi, sb) : 0u;
            nj = oj < nj ? min(nj - oj, sb) : 0u;
            i0 += oi;
            j0 += oj;
        }
        const bool same_tile = e.bi == e.bj;
        uint64_t local = 0;
        for (uint32_t p = wave; p < ni * nj; p += blockDim.x / 64) {
            const uint32_t a = i0 + p / nj, b = j0 + p % nj;
            if (same_tile && b <= a) continue;
            const uint32_t i = D.perm[a], j = D.perm[b];
            uint32_t cnt = 0;
            int cell = STORM_HIP_CELL_EMPTY;
            if (!span_disjoint(D, i, j)) {
                const bool i_dense = meta(D, i, 0) >= meta(D, j, 0);
                const TileMeta md = row_tile(D, i_dense ? i : j), ms = row_tile(D, i_dense ? j : i);
                if (first == kPerPairCell) {
                    const uint32_t elig = eligible_mask(md, ms, D.reps) & 0x3FFu;
                    double c;
                    cell = select_pairing(M, md, ms, &c);
                    if (cell == STORM_HIP_CELL_EMPTY || !(elig >> cell & 1))
                        cell = cheapest_eligible(M, md, ms, elig);
                } else {
                    cell = predict(M, second, md, ms) < predict(M, first, md, ms) ? second : first;
                }
                cnt = wave_sum_u32(pair_partial(cell, D, M, i, j, lane, 64));
            }
            if (lane == 0) {
                atomicAdd(&s_pairs[cell], 1ull);
                local += cnt;
                emit_pair(D, i, j, cnt);
            }
        }
        local = wave_sum_u64(local);
        if (lane == 0 && local) atomicAdd(D.sum, (unsigned long long)local);
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint64_t now = wall_clock64();
            if (tile_cycles) tile_cycles[t] = now - t0;
            if (D.finish) atomicMax(D.finish, (unsigned long long)now);
        }
    }
    // Pairs per cell run; span-test skips count as EMPTY and as skipped.
    if (threadIdx.x < kNumCells && s_pairs[threadIdx.x]) {
        atomicAdd(&ctr->cell_pairs[threadIdx.x], s_pairs[threadIdx.x]);
        if (threadIdx.x == STORM_HIP_CELL_EMPTY) atomicAdd(&ctr->skipped, s_pairs[threadIdx.x]);
    }
}

int launch_refine(hipStream_t s, uint32_t grid, const DevCorpus* Ds, const DevModel& M,
                  const QItem* queue, DevCounters* ctr, unsigned long long* tile_cycles) {
    hipLaunchKernelGGL(refine_kernel, dim3(grid), dim3(256), 0, s, Ds, M, queue, ctr, tile_cycles);
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

uint32_t refine_blocks_per_cu() {
    static std::atomic<int> cache{0};
    int v = cache.load(std::memory_order_relaxed);
    if (v > 0) return (uint32_t)v;
    int nb = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, refine_kernel, 256, 0);
    v = (e == hipSuccess && nb > 0) ? nb : 4;
    cache.store(v, std::memory_order_relaxed);
    return (uint32_t)v;
}

// Workgroups of a cell kernel resident on one CU at once (registers, LDS and
// waves permitting), queried once per cell.
uint32_t cell_blocks_per_cu(int cell) {
    static std::atomic<int> cache[kNumCells];
    if (cell < 0 || cell >= kNumCells) return 1;
    int v = cache[cell].load(std::memory_order_relaxed);
    if (v > 0) return (uint32_t)v;
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
#define SH_OCC(C) \
    case C: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, cell_kernel<C>, 256, 0); break;
    switch (cell) {
        SH_OCC(STORM_HIP_CELL_BS) SH_OCC(STORM_HIP_CELL_BR) SH_OCC(STORM_HIP_CELL_BW)
        SH_OCC(STORM_HIP_CELL_SS) SH_OCC(STORM_HIP_CELL_SR) SH_OCC(STORM_HIP_CELL_SW)
        SH_OCC(STORM_HIP_CELL_RR) SH_OCC(STORM_HIP_CELL_RW) SH_OCC(STORM_HIP_CELL_WW)
        SH_OCC(STORM_HIP_CELL_EMPTY) SH_OCC(STORM_HIP_CELL_CB)
        default: break;
    }
#undef SH_OCC
    v = (e == hipSuccess && nb > 0) ? nb : 4;
    cache[cell].store(v, std::memory_order_relaxed);
    return (uint32_t)v;
}

int launch_cell(int cell, hipStream_t s, uint32_t grid, const DevCorpus* Ds, const DevModel& M,
                const QItem* queue, DevCounters* ctr, unsigned long long* tile_cycles) {
#define SH_CELL(C)                                                                          \
    case C:                                                                                 \
        hipLaunchKernelGGL(cell_kernel<C>, dim3(grid), dim3(256), 0, s, Ds, M, queue, ctr,  \
                           tile_cycles);                                                     \
        break;
    switch (cell) {
        SH_CELL(STORM_HIP_CELL_BS)
        SH_CELL(STORM_HIP_CELL_BR)
        SH_CELL(STORM_HIP_CELL_BW)
        SH_CELL(STORM_HIP_CELL_SS)
        SH_CELL(STORM_HIP_CELL_SR)
        SH_CELL(STORM_HIP_CELL_SW)
        SH_CELL(STORM_HIP_CELL_RR)
        SH_CELL(STORM_HIP_CELL_RW)
        SH_CELL(STORM_HIP_CELL_WW)
        SH_CELL(STORM_HIP_CELL_EMPTY)
        SH_CELL(STORM_HIP_CELL_CB)
        default:
            set_error("launch_cell: cell %d has no sparse kernel", cell);
            return STORM_HIP_ERR_INVALID;
    }
#undef SH_CELL
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

}  // namespace stormhip
