// Device-side view of a general (mixed-representation) corpus and the
// structures the selection and cell kernels share.
#pragma once

#include <cstdint>

#include "device_common.h"
#include "internal.h"

namespace stormhip {

constexpr uint32_t kNoBitmap = 0xFFFFFFFFu;
constexpr int kNumCells = STORM_HIP_NCELL;   // BB..WW, EMPTY, CB

// One reported pair of a thresholded call (layout of storm_hip_pair).
struct PairHit { uint32_t i, j, count; };

// Row arenas on the device. Row ids here are ORIGINAL ids; `perm` maps a
// density-sorted position to its original id.
struct DevCorpus {
    uint32_t n, universe, nw, ld;
    const uint64_t* list_off; const uint32_t* list;
    const uint64_t* run_off;  const uint32_t* run_start; const uint32_t* run_end;
    const uint64_t* ewah_off; const uint64_t* ewah;
    // EWAH marker table: for marker m of row r (m in [mk_off[r], mk_off[r+1])):
    // mk_word = decoded index of its first word, mk_pos = its offset in the row's
    // EWAH buffer.
    const uint64_t* mk_off;   const uint32_t* mk_word; const uint32_t* mk_pos;
    const uint64_t* comp_off; const uint32_t* comp;
    const uint32_t* bm_row;   const uint64_t* bitmap;   // bitmap row index, kNoBitmap if none
    // rank9 and zone map of a bitmap row (kNoBitmap in the row map: none):
    // rank + rank_row[r] * rank_ld, occ + occ_row[r] * occ_ld.
    const uint32_t* rank_row; const uint64_t* rank; uint32_t rank_ld;
    const uint32_t* occ_row;  const uint64_t* occ;  uint32_t occ_ld;
    const uint32_t* meta;     // 8 per original row (RowMeta + has_rank, has_occ)
    const uint32_t* perm;     // sorted position -> original row id
    // Per position of perm's order: the row's cardinality, first and last
    // position (null when perm is the identity order: read meta instead).
    const uint4* pspan;
    // Bit-sliced bitmaps of 64-row groups of the density order, eight groups
    // (512 rows) per slab: with s = slice_of[g] (kNoBitmap: none), word
    // slice[((s / 8) * ld * 64 + p) * 8 + s % 8] has bit k set when sorted
    // row 64g + k holds position p; groups 8q..8q+7 share slab q, lane g % 8.
    const uint64_t* slice;    const uint32_t* slice_of;
    // Coarse zone maps (null: none): row r's zone_ld words at zone + r *
    // zone_ld, bit b for positions p with p >> zone_shift == b.
    const uint64_t* zone;     uint32_t zone_ld, zone_shift;
    // Per-call state of this corpus inside a (possibly multi-corpus) call.
    const struct TileMeta* tm;        // its row tiles' metadata
    uint64_t* tzone;                  // per row tile: OR of its rows' zone maps
    const uint32_t* rowmap;           // sorted position -> bitmap row
    uint32_t* tri;                    // per-pair output or null
    // 1: tri is packed over density positions (inv maps a row to its
    // position); 0: over original rows.
    uint32_t tri_pos;
    unsigned long long* sum;          // its sum
    unsigned long long* finish;       // wall clock of its last finished tile (max)
    uint32_t tile;                    // rows per tile
    // 256-row B x B tiles are queued as this many K slices (1: whole tiles);
    // the slices add into the corpus's triangle, zeroed before the call.
    uint32_t bb_ksplit;
    // Thresholded calls (hits non-null, tri null): pairs whose Jaccard
    // |A n B| / |A u B| is at least thr are appended to hits (n_hits counts
    // every such pair, including those past hit_cap).
    double thr;
    PairHit* hits;
    unsigned long long* n_hits;
    uint64_t hit_cap;
    // Band and cross calls (emit_kind 1 / 2, tri null): counts go to emit_buf
    // by position in the call's order (inv: original row -> position; also
    // the map of a position-packed triangle):
    //   1  band of one tile row:  emit_buf[(lo % tile) * n + hi]
    //   2  cross rows [0, split) x [split, n): emit_buf[lo * emit_cols + hi - split]
    uint32_t emit_kind, emit_cols, emit_split;
    uint32_t* emit_buf;
    const uint32_t* inv;
    // Shadow instance of a probe call: the same corpus, no output (tri, hits
    // and emit off, sum into a scratch counter, pair counts not recorded).
    uint32_t shadow;
    uint16_t* probe;                  // per-tile probe state of a probe call
    uint32_t reps;                    // representations held (kRep* bits)
};

// Per-pair output of a general call: the packed triangle, or the pair itself
// when a thresholded call's Jaccard test passes -- the test of the
// thresholded benchmark (bench/bench_threshold.cpp jac_ok): u = |A| + |B| -
// |A n B| in double, reported when |A n B| > 0 and |A n B| / u >= thr.
__device__ inline void emit_pair(const DevCorpus& D, uint32_t i, uint32_t j, uint32_t v) {
    if (D.tri) {
        if (D.tri_pos) { i = D.inv[i]; j = D.inv[j]; }
        D.tri[i < j ? tri_index(i, j, D.n) : tri_index(j, i, D.n)] = v;
        return;
    }
    if (D.emit_kind) {
        const uint32_t a = D.inv[i], b = D.inv[j];
        const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        if (hi >= D.n) return;   // padding positions carry no pair
        if (D.emit_kind == 1)
            D.emit_buf[(uint64_t)(lo % D.tile) * D.n + hi] = v;
        else if (lo < D.emit_split && hi >= D.emit_split)
            D.emit_buf[(uint64_t)lo * D.emit_cols + (hi - D.emit_split)] = v;
        return;
    }
    if (D.hits && v) {
        const double ci = D.meta[(uint64_t)i * 8], cj = D.meta[(uint64_t)j * 8];
        if ((double)v / (ci + cj - (double)v) >= D.thr) {
            // Global-address-space accesses: a flat store could alias a
            // kernel's private arrays and pin them in scratch.
            typedef __attribute__((address_space(1))) unsigned long long g_u64;
            typedef __attribute__((address_space(1))) uint32_t g_u32;
            const unsigned long long k = __hip_atomic_fetch_add(
                (g_u64*)D.n_hits, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (k < D.hit_cap) {
                g_u32* h = (g_u32*)(D.hits + k);
                h[0] = i < j ? i : j;
                h[1] = i < j ? j : i;
                h[2] = v;
            }
        }
    }
}

// emit_pair for a pair whose density positions a and b are at hand (rows
// i = perm[a], j = perm[b]): a position-packed triangle takes them as they are.
__device__ inline void emit_at(const DevCorpus& D, uint32_t a, uint32_t b, uint32_t i, uint32_t j,
                               uint32_t v) {
    if (D.tri && D.tri_pos) {
        D.tri[a < b ? tri_index(a, b, D.n) : tri_index(b, a, D.n)] = v;
        return;
    }
    emit_pair(D, i, j, v);
}

// One queued tile pair: corpus index, tile row, tile column, and `sub`: 0 for
// the whole tile, else one block of it: (sub & 0xFFFF) rows per side, the
// ((sub >> 16) & 0xFF)-th such block of the tile's rows and the (sub >> 24)-th
// of its columns. Per-pair cells take 256-row tiles as 64 x 64 blocks (more
// work items than CUs several times over); the selector's sampling runs time
// one block of median-density rows and scale by the pair ratio.
struct QItem { uint32_t inst, bi, bj, sub; };

__host__ __device__ inline uint32_t make_sub(uint32_t rows, uint32_t bi_blk, uint32_t bj_blk) {
    return rows | (bi_blk << 16) | (bj_blk << 24);
}

// Rows per block when a tile is split for the per-pair cells.
constexpr uint32_t kSplitRows = 64;

// B x B items: sub = slice | (k_split << 16) names K slice `slice` of k_split
// (sub 0: the whole tile). A 256-row tile of a corpus with n_words words is
// cut into about n_words / kBBSliceWords slices (at most kBBMaxSlices), so
// that a wide universe does not leave one workgroup per tile running for
// seconds while the rest of the chip idles.
constexpr uint32_t kBBSliceWords = 32768, kBBMaxSlices = 64;
__host__ __device__ inline uint32_t bb_slices_for(uint32_t n_words) {
    const uint32_t k = (n_words + kBBSliceWords - 1) / kBBSliceWords;
    return k < 1 ? 1 : (k > kBBMaxSlices ? kBBMaxSlices : k);
}
// A bound on the words one of ks K slices of a 256-row B x B tile covers
// (whole stages of up to 16 words each): the f32 counts of an fp4 item are
// exact while this stays below 2^18, whatever the universe, because the
// slices add into the output as integers.
__host__ __device__ inline uint32_t bb_item_words(uint32_t n_words, uint32_t ks) {
    return (n_words + ks - 1) / (ks ? ks : 1) + 32;
}

// A block index of kWholeSide in a sub means "the whole side of the tile":
// B x S items are stripes, kSplitRows rows of the probed (list) side against
// every row of the other side, whose bitmaps are read through their slices.
// On a diagonal tile a stripe keeps only the pairs above the diagonal.
constexpr uint32_t kWholeSide = 0xFF;
// kWideSide in place of the column block: a wide stripe, kSplitRows list rows
// of tile bi against both bitmap tiles bj and bj + 1 (bj even, both full and
// sliced, bi < bj): one 64-byte slice load per list element answers for 512
// rows. The select kernel queues one when both tile pairs chose B x S.
constexpr uint32_t kWideSide = 0xFE;

// Bit-sliced bitmaps come in slabs of kSlabGroups 64-row groups (the two
// 256-row tiles 2k, 2k+1): per universe position the slab keeps its groups'
// masks side by side, kSlabGroups x 8 bytes.
constexpr uint32_t kSlabGroups = 8;

// Thresholded calls (the corpora carry a Jaccard threshold and pair buffer,
// their tri is null) use the B x B instances that report pairs; the matrix-core
// one takes thr_buf, grid blocks x 256 x 256 u32 of per-block count buffers.
int launch_bb_valu_queue(hipStream_t s, int bm, const DevCorpus* Ds, const QItem* queue,
                         const unsigned int* qcount, unsigned int* qhead, uint32_t grid,
                         unsigned long long* tile_ticks, bool thresholded = false);
int launch_bb_mfma_queue(bool fp4, const MfmaConfig& cfg, hipStream_t s, const DevCorpus* Ds,
                         const QItem* queue, const unsigned int* qcount, unsigned int* qhead,
                         uint32_t grid, unsigned long long* tile_ticks,
                         uint32_t* thr_buf = nullptr);
// acc_words: the most words one work item accumulates in its registers (a K
// slice of a wide tile; 0: n_words) -- fp4's f32 counts are exact below 2^18.
int bb_mfma_check(bool fp4, uint32_t stage_words, uint32_t n_words, uint32_t ld,
                  uint32_t acc_words = 0);
// Workgroups resident per CU: a cell kernel; the B x B kernel of a tile width
// (256 = matrix cores with this launch shape, else VALU).
uint32_t cell_blocks_per_cu(int cell);
uint32_t bb_blocks_per_cu(uint32_t tile, const MfmaConfig& cfg);

// Aggregate metadata of one row tile (density-sorted rows [lo, hi)). The first
// block is tile_meta of kernels/storm_allpairs.cpp:281-312; the rest are
// eligibility flags and spread statistics for the kernels and the selector.
struct TileMeta {
    uint32_t card, runs, nnz, nw, first, last;   // means (integer), min first, max last
    uint32_t has_rank, has_occ;                  // AND over rows
    uint32_t all_bitmap, all_comp, any_empty;    // AND / AND / OR over rows
    uint32_t max_card, min_card, max_runs, max_ewah, rows;
    double spread;                               // max / mean cardinality
};

// Buckets shared by the zone maps of row tiles bi and bj (tz: one map of
// zone_ld words per tile): 0 proves every pair of the tile pair disjoint.
__device__ inline uint32_t tile_zone_shared(const uint64_t* __restrict__ tz, uint32_t zone_ld,
                                            uint32_t bi, uint32_t bj) {
    const uint64_t* a = tz + (uint64_t)bi * zone_ld;
    const uint64_t* b = tz + (uint64_t)bj * zone_ld;
    uint32_t c = 0;
    for (uint32_t w = 0; w < zone_ld; ++w) c += __popcll(a[w] & b[w]);
    return c;
}

// B x S probes the lists of side A (else B) against the other side's bitmaps:
// A when B is all bitmaps and A is the sparser side (or not all bitmaps).
__host__ __device__ inline bool bs_probe_a(const TileMeta& a, const TileMeta& b) {
    return b.all_bitmap && (!a.all_bitmap || a.card <= b.card);
}

// The stripes serve a B x S item when the probed side's mean cardinality is
// at most 64x the sparser side's (probing the denser side costs its elements
// once per stripe; the per-pair cell probes the sparser side's once per pair).
__host__ __device__ inline bool bs_stripes_pay(uint32_t probe_card, const TileMeta& a,
                                               const TileMeta& b) {
    return (double)probe_card <= 64.0 * (double)(a.card < b.card ? a.card : b.card) + 64.0;
}

// Tile pairs (bi, bj) and (bi, bj + 1) that both chose B x S run as wide
// stripes: bj even, bi < bj, both bitmap tiles full and all bitmaps (so their
// eight groups have slices), the list side bi probed against each.
__host__ __device__ inline bool bs_wide_ok(uint32_t bi, uint32_t bj, uint32_t tile,
                                           const TileMeta& a, const TileMeta& b0,
                                           const TileMeta& b1) {
    return tile == 256 && (bj & 1) == 0 && bi < bj && b0.rows == 256 && b1.rows == 256 &&
           b0.all_bitmap && b1.all_bitmap && bs_probe_a(a, b0) && bs_probe_a(a, b1) &&
           bs_stripes_pay(a.card, a, b0) && bs_stripes_pay(a.card, a, b1);
}

// Representations held for every row of a corpus (reps bits).
enum : uint32_t { kRepRuns = 1, kRepEwah = 2, kRepComp = 4, kRepLists = 8 };

// Cells a tile pair can run given the representations resident on the device
// (bit per STORM_HIP_CELL_*): bitmaps per row (the tiles' all_bitmap), the
// other representations per corpus (reps). C x B takes the complement as a
// list or, from the runs, as the gaps between them.
__host__ __device__ inline uint32_t eligible_mask(const TileMeta& a, const TileMeta& b,
                                                  uint32_t reps) {
    const bool runs = reps & kRepRuns, ewah = reps & kRepEwah, comp = reps & kRepComp;
    const bool lists = reps & kRepLists;
    const bool bb = a.all_bitmap && b.all_bitmap, b1 = a.all_bitmap || b.all_bitmap;
    uint32_t e = 0;
    if (bb) e |= 1u << STORM_HIP_CELL_BB;
    if (b1 && lists) e |= 1u << STORM_HIP_CELL_BS;
    if (b1 && runs) e |= 1u << STORM_HIP_CELL_BR;
    if (b1 && ewah) e |= 1u << STORM_HIP_CELL_BW;
    if (lists) e |= 1u << STORM_HIP_CELL_SS;
    if (runs && lists) e |= 1u << STORM_HIP_CELL_SR;
    if (runs) e |= 1u << STORM_HIP_CELL_RR;
    if (ewah && lists) e |= 1u << STORM_HIP_CELL_SW;
    if (ewah) e |= 1u << STORM_HIP_CELL_WW;
    if (runs && ewah) e |= 1u << STORM_HIP_CELL_RW;
    if ((comp || runs) && ((a.all_comp && b.all_bitmap) || (b.all_comp && a.all_bitmap)))
        e |= 1u << STORM_HIP_CELL_CB;
    return e;
}

// Sound tile-level emptiness: every pair of the tile fails the span test.
__host__ __device__ inline bool tile_empty(const TileMeta& a, const TileMeta& b) {
    return a.max_card == 0 || b.max_card == 0 || a.last < b.first || b.last < a.first;
}

// Largest min(|A|, |B|) / max(|A|, |B|) over the tile's pairs: an upper bound
// on every pair's Jaccard (|A n B| <= min, |A u B| >= max). Computed as the
// same double division as a pair's Jaccard in the equality case, so a tile
// whose bound is below the threshold holds no reportable pair.
__host__ __device__ inline double length_bound(const TileMeta& a, const TileMeta& b) {
    if (a.max_card >= b.min_card && b.max_card >= a.min_card) return 1.0;
    return a.max_card < b.min_card ? (double)a.max_card / (double)b.min_card
                                   : (double)b.max_card / (double)a.min_card;
}

// Device cost-model knobs passed by value.
struct DevModel {
    double ns_per_unit[10];
    double ns_fixed, ns_per_occ_word, ns_probe_hot, ns_probe_cold;
    double probe_scale, br_margin, rr_scale;
    // 1: B x B priced as the reference's zone-mapped kernel where both sides
    // carry a zone map; 0: always Theta(m) words, as the device B x B runs.
    double bb_zonemap;
};

// Internal policy of exploration calls: each non-empty tile pair runs a cell
// drawn (seeded) from the cells the learned selector may rank on it.
constexpr uint32_t kPolicyExplore = 16;
// The two selection passes of a PROBE call: pass 1 runs every tile that has
// a single candidate and, on the sampled tiles, both of its two best
// candidates (the second on the corpus's shadow instance, which writes no
// output); pass 2 commits the pending tiles from the measured times.
constexpr uint32_t kPolicyProbe1 = 17, kPolicyProbe2 = 18;
// Probe classes: (first, second, predicted-ratio bucket) of a tile's two best
// candidates; sampled tiles of a class decide for its pending tiles.
constexpr uint32_t kProbeBuckets = 4;
constexpr uint32_t kProbeClasses = 10 * 10 * kProbeBuckets;
// Every class samples at least this many tiles, then one tile in 2^24 /
// kProbeShare more.
constexpr uint32_t kProbeMinSamples = 4;
constexpr uint32_t kProbeShare = 1u << 19;   // 1 / 32
// Per-tile probe state (u16): first | second << 4 | bucket << 8, plus
// kProbeSampled or kProbePending; 0 for a tile settled in pass 1.
constexpr uint16_t kProbeSampled = 1u << 12, kProbePending = 1u << 13;

struct ProbeAcc {
    unsigned long long ns[2][kProbeClasses];   // CU-time of first / second
    unsigned int tiles[kProbeClasses];         // tiles seen per class
    unsigned long long sampled, switched;      // tiles sampled; committed to second
};

// Workgroups per CU of each cell's kernel in this call (B x B: the launched
// form), to turn an item's wall time into CU-time.
struct ProbeBpc { float v[kNumCells]; };

// Refine (kernels/storm_allpairs.cpp:563-588): a tile whose rows spread by at
// least kRefineHet (max / mean cardinality) and whose two best candidates
// are within kRefineRatio of each other lets each pair pick between them.
constexpr double kRefineHet = 8.0, kRefineRatio = 3.0;

// Learned selection ranks only cells predicted (analytically) within this
// factor of the cheapest eligible cell on the tile.
constexpr double kLearnedPrune = 32.0;
// ... and replaces the analytic choice when it predicts its own pick faster
// by more than this many log2 units (a trust region around the analytic
// model). At 0.5 (a predicted gain of sqrt(2)) the network's errors on near
// ties stay out: on the skewed and run-heavy workloads, margins 0 / 0.25 /
// 0.5 / 1.0 gave 1.00 / 1.09 / 1.13 / 1.14x and 0.91 / 1.03 / 0.99 / 0.97x
// over the better analytic model (about 4% run-to-run noise).
constexpr float kLearnedMargin = 0.5f;
__host__ __device__ uint32_t learned_candidates(const DevModel& m, const TileMeta& a,
                                                const TileMeta& b, uint32_t elig);

// B x B tiles of corpora whose counts can exceed fp4's exact range (2^18
// words or more) go to a queue of their own, run by the int8 matrix-core
// kernel, so the other corpora of a batch keep the fp4 one.
constexpr int kQueueBBWide = kNumCells;
// Refined tiles: each pair runs the cheaper (by its own rows' metadata) of
// the tile's two candidates, named in the item's inst (kRefineShift).
constexpr int kQueueRefine = kNumCells + 1;
constexpr int kNumQueues = kNumCells + 2;
constexpr uint32_t kRefineShift = 24;
// Candidates of a per-pair item (PerPair / Oracle): every cell, per pair.
constexpr int kPerPairCell = 15;

// Counters written by the selection and cell kernels.
struct DevCounters {
    unsigned long long cell_pairs[kNumCells];
    unsigned long long cell_tiles[kNumCells];
    unsigned long long skipped;
    unsigned long long fallback_tiles;
    unsigned int qcount[kNumQueues];   // items queued per queue
    unsigned int qhead[kNumQueues];    // next item to take, per queue
};

// Probe: fold pass 1's timed items into their classes, then empty the
// queues for pass 2.
int launch_probe_reduce(hipStream_t s, const DevCorpus* Ds, uint32_t n_inst, const QItem* queue,
                        uint64_t qstride, const DevCounters* ctr,
                        const unsigned long long* ticks, const ProbeBpc& bpc, ProbeAcc* acc);

}  // namespace stormhip
