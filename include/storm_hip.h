/*
 * storm_hip.h -- C ABI of the MI355X (gfx950) exact all-pairs intersection
 * cardinality engine.
 *
 * Plain C: opaque handles, plain pointers and sizes, integer status codes,
 * no exceptions and no device-framework types cross this boundary.
 *
 * Each entry point names the StormBitmaps interface it stands in for
 * (paths relative to the StormBitmaps source tree).
 */
#ifndef STORM_HIP_H_
#define STORM_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STORM_HIP_ABI_VERSION 3

/* Status codes. Negative = failure; the reference's C ABI uses the same
 * convention (storm.cpp:1160-1219: NULL handle -1, NULL values -2, bad_alloc -4). */
enum {
    STORM_HIP_OK              = 0,
    STORM_HIP_ERR_INVALID     = -1,  /* NULL handle or argument out of contract */
    STORM_HIP_ERR_NULL_VALUES = -2,  /* NULL data pointer with a nonzero size   */
    STORM_HIP_ERR_DEVICE      = -3,  /* HIP runtime or launch failure           */
    STORM_HIP_ERR_NOMEM       = -4,  /* device or host allocation failed        */
    STORM_HIP_ERR_UNSUPPORTED = -5,  /* combination not available               */
    STORM_HIP_ERR_RANGE       = -6   /* size beyond what the engine supports    */
};

typedef struct storm_hip_ctx    storm_hip_ctx;
typedef struct storm_hip_corpus storm_hip_corpus;

const char* storm_hip_strerror(int status);
/* Library version string, e.g. "storm_hip 0.1 gfx950 abi 1". */
const char* storm_hip_version(void);
/* Number of visible HIP devices (0 when none). */
int storm_hip_device_count(int* n_devices);
/* Last error message recorded on this thread (never NULL). */
const char* storm_hip_last_error(void);

/* One engine instance bound to one device and one HIP stream. One in-flight
 * call per context; calls on one context are stream-ordered. */
int  storm_hip_ctx_create(int device, uint32_t flags, storm_hip_ctx** out);
void storm_hip_ctx_free(storm_hip_ctx* ctx);
int  storm_hip_sync(storm_hip_ctx* ctx);

/* ---- corpora -------------------------------------------------------------
 *
 * Dense contiguous corpus: the device form of STORM_contiguous_t
 * (storm.cpp:1113-1157): n_rows bitmaps of n_words 64-bit words each, row
 * major. Replaces STORM_contig_new + STORM_contig_add (storm.cpp:1159-1220)
 * for the bulk-upload case. The words are copied; the caller keeps ownership.
 */
int storm_hip_corpus_upload_contig(storm_hip_ctx* ctx, const uint64_t* words,
                                   uint32_t n_rows, uint32_t n_words,
                                   storm_hip_corpus** out);
/* Same corpus shape, generated on the device: bit b of row r is set iff the
 * (r*n_bits+b)-th draw u of the splitmix64 stream Rng(seed)
 * (kernels/storm_gen.h:84-102), taken as unit() = (next()>>11)*2^-53,
 * satisfies u < p. Bits past n_bits in the last word are zero. */
int storm_hip_corpus_gen_bernoulli(storm_hip_ctx* ctx, uint32_t n_rows, uint32_t n_bits,
                                   double p, uint64_t seed, storm_hip_corpus** out);
/* Copy rows [row0, row0+n_rows) of a dense corpus back to the host. */
int storm_hip_corpus_download_words(storm_hip_ctx* ctx, const storm_hip_corpus* c,
                                    uint32_t row0, uint32_t n_rows, uint64_t* out);

/* General corpus: per-row arenas in the storm_repr byte layouts
 * (kernels/storm_repr.h:112-276), the layouts build_row writes
 * (kernels/storm_repr.cpp:102-207). All offsets are element offsets into the
 * matching arena, n_rows+1 entries each (row r spans [off[r], off[r+1]); the
 * first offset may be non-zero). An arena whose offsets are NULL is absent.
 *   lists        sorted distinct u32 positions          (ListView)
 *   runs         [start,end) u32 SoA, maximal           (RunView)
 *   ewah         EWAH-64 markers + literals             (EwahView)
 *   bitmap       n_words u64 per row, or an empty span  (BitmapView::w)
 *   rank         rank9, 2*(ceil(n_words/8)+1) u64       (BitmapView::rank)
 *   occ          zone map, 1 bit per 8 words            (BitmapView::occ)
 *   comp lists   complement positions, rows with 2n > universe (ComplementView)
 *   meta         8 u32 per row: cardinality, n_runs, n_words, n_nonzero_w,
 *                first_set, last_set, has_rank, has_occ  (RowMeta)
 * Every row needs a list, runs or a bitmap; lists are optional, so rows held
 * only as runs or only as bitmaps can be uploaded (cells that read lists are
 * then not eligible). */
typedef struct {
    uint32_t        n_rows;
    uint32_t        universe;
    const uint64_t* list_off;   const uint32_t* list;
    const uint64_t* run_off;    const uint32_t* run_start;  const uint32_t* run_end;
    const uint64_t* ewah_off;   const uint64_t* ewah;
    const uint64_t* bitmap_off; const uint64_t* bitmap;
    const uint64_t* rank_off;   const uint64_t* rank;
    const uint64_t* occ_off;    const uint64_t* occ;
    const uint64_t* comp_off;   const uint32_t* comp;
    const uint32_t* meta;       /* 8 per row */
} storm_hip_arena;

/* Upload from arenas, each taken as given after a host check against the row
 * it describes (sorted distinct lists inside the universe, maximal runs, EWAH
 * decoding to the row, complements, bitmaps of n_words words with no bit past
 * the universe, rank9 / zone map equal to build_rank / build_occ of the
 * bitmap, RowMeta equal to the row's; a row without a bitmap may carry
 * build_row's sparse-only meta, n_nonzero_w = n_runs and no index flags).
 * A bitmap is placed on the device for exactly the rows whose bitmap span is
 * non-empty; rank9 and zone maps for those of them whose meta has_rank /
 * has_occ is set (built on the device, the same bytes). Violations are
 * STORM_HIP_ERR_INVALID with the row and the reason in storm_hip_last_error. */
int  storm_hip_corpus_upload(storm_hip_ctx* ctx, const storm_hip_arena* arena,
                             storm_hip_corpus** out);

/* Build every representation from sorted, distinct per-row positions (CSR:
 * row r is pos[offs[r]..offs[r+1])), with the semantics of build_row
 * (kernels/storm_repr.cpp:102-207) including its rank/occ/complement gates in
 * RowMeta, and upload. The reference equivalent is building std::vector<Row>
 * with build_row before allpairs_sum. flags: bits 0-1 bitmap placement
 * (0: rows with cardinality*64 >= universe, 1: every row, 2: none), plus the
 * STORM_HIP_BUILD_* representation bits. The lists are uploaded and the
 * representations built on the device, one wavefront per row; with
 * STORM_HIP_BUILD_ON_HOST they are built by host threads and uploaded (same
 * arenas byte for byte). STORM_HIP_BUILD_SLICES adds a bit-sliced copy of
 * the bitmaps of every 64 density-ordered rows that all hold one (one 64-bit
 * row mask per position; as much device memory as those bitmaps), which lets
 * B x S count a 64-row block against a list with one load per list element;
 * it has no counterpart in the reference's representations.
 * STORM_HIP_BUILD_ZONES adds a coarse positional zone map per row (the Occ of
 * bench/bench_tile.cpp:52-62 at 16,384 buckets: bucket = position >> shift,
 * shift the smallest with 2^shift >= ceil(universe / 16384), at most 257
 * words per row), built on the device from whichever representation the row
 * holds. Calls then settle as empty every tile pair whose tiles share no
 * bucket (the OR of their rows' maps; the empty-tile skip of
 * bench_tile.cpp:807-827) and, where a tile pair shares under half of the
 * buckets, skip in the per-pair cells every pair whose rows share none (the
 * gated per-pair filter of bench_tile.cpp:180-190). Counts are unchanged;
 * pairs skipped this way count as skipped, like the span test's. */
enum {
    STORM_HIP_BITMAP_DENSE = 0, STORM_HIP_BITMAP_ALL = 1, STORM_HIP_BITMAP_NONE = 2,
    STORM_HIP_BUILD_RUNS = 4, STORM_HIP_BUILD_EWAH = 8, STORM_HIP_BUILD_COMP = 16,
    STORM_HIP_BUILD_ALL = 4 | 8 | 16,
    STORM_HIP_BUILD_ON_HOST = 32,
    STORM_HIP_BUILD_SLICES = 64,
    STORM_HIP_BUILD_ZONES = 128
};
/* Seeded synthetic corpora as host CSR lists (storm::generate's role,
 * kernels/storm_gen.cpp:173-232, with the spectra the benchmark workloads
 * need). Row r draws from its own splitmix64 stream seeded by draw r of
 * Rng(seed) (kernels/storm_gen.h:84-102).
 *   ZIPF       cardinality Zipf(s=p1) on [1, p2], positions uniform without
 *              replacement
 *   RUNS       Poisson(p1) runs, geometric lengths of mean p2, uniform starts,
 *              plus round(p3 * run bits) uniform singletons
 *   BERNOULLI  every bit i.i.d. Bernoulli(p1) (geometric skips)
 * *offs (n_rows+1) and *pos are allocated by the library; release each with
 * storm_hip_gen_free. */
enum { STORM_HIP_GEN_ZIPF = 0, STORM_HIP_GEN_RUNS = 1, STORM_HIP_GEN_BERNOULLI = 2 };
int  storm_hip_gen_corpus(uint32_t kind, uint32_t n_rows, uint32_t universe, double p1,
                          double p2, double p3, uint64_t seed, uint64_t** offs, uint32_t** pos,
                          uint64_t* total);
void storm_hip_gen_free(void* p);

int  storm_hip_corpus_from_lists(storm_hip_ctx* ctx, const uint64_t* offs, const uint32_t* pos,
                                 uint32_t n_rows, uint32_t universe, uint32_t flags,
                                 storm_hip_corpus** out);
/* As storm_hip_corpus_from_lists for lists already in device memory (d_offs:
 * n_rows + 1 u64 offsets, d_pos: u32 positions); the corpus keeps copies and
 * builds on the device (STORM_HIP_BUILD_ON_HOST is ignored). */
int  storm_hip_corpus_from_device_lists(storm_hip_ctx* ctx, const uint64_t* d_offs,
                                        const uint32_t* d_pos, uint32_t n_rows,
                                        uint32_t universe, uint32_t flags,
                                        storm_hip_corpus** out);
/* The seeded workloads generated on the device, each row in the
 * representation it is stored in (no host lists, no upload):
 *   BERNOULLI, p1 >= 1/32  every row a bitmap only: bit b of row r set iff
 *                          draw r*universe+b of Rng(seed) is below p1 (as
 *                          storm_hip_corpus_gen_bernoulli)
 *   BERNOULLI, p1 < 1/32   lists, the rows of storm_hip_gen_corpus
 *                          (geometric skips on each row's stream), then built
 *                          like storm_hip_corpus_from_lists with `flags`
 *   RUNS                   maximal runs only, the rows of storm_hip_gen_corpus
 *                          (Poisson(p1) runs of mean length p2 + p3 singletons);
 *                          bitmaps placed by flags bits 0-1
 * The same position sets as the host generator, bit for bit (both draw from
 * the same row streams through the same deterministic log and exp).
 * ZIPF is STORM_HIP_ERR_UNSUPPORTED here. */
int  storm_hip_corpus_gen(storm_hip_ctx* ctx, uint32_t kind, uint32_t n_rows, uint32_t universe,
                          double p1, double p2, double p3, uint64_t seed, uint32_t flags,
                          storm_hip_corpus** out);
/* Column closed forms of a corpus (no counterpart in the reference): with c_b
 * the number of rows holding position b, *total = sum_b C(c_b, 2) -- the sum
 * over pairs i < j of |Xi n Xj| -- and row_sums[i] (optional, n_rows) =
 * sum_{b in Xi} (c_b - 1) = sum_{j != i} |Xi n Xj|. Computed from the rows'
 * own representation (lists, else runs, else bitmaps) without any cell: an
 * exact check of an all-pairs result at any size. */
int  storm_hip_corpus_closed_form(storm_hip_ctx* ctx, const storm_hip_corpus* c, uint64_t* total,
                                  uint64_t* row_sums);

/* STORMBIN ingest. Format (tools/sets2bin.py:23-25): b"STORMBIN", then u32
 * version, n_rows, n_bits, pad, then per row a u32 count followed by count
 * ascending u32 positions. As the reference loader (load_stormbin,
 * bench/bench_allpairs.cpp:44-72): rows are read to the end of the file or the
 * first truncated record; every `stride`-th row (stride 0 means 1) is kept
 * unless it is empty; at most max_rows rows are kept (0: no limit); the
 * universe is n_bits; fewer than 2 kept rows is STORM_HIP_ERR_INVALID.
 * storm_hip_stormbin_read returns host CSR lists (release with
 * storm_hip_gen_free); storm_hip_corpus_from_stormbin uploads the row records
 * as they lie in the file, gathers them into the list on the device and
 * builds every representation there (flags as storm_hip_corpus_from_lists).
 * Corpus row k is the k-th kept row. */
int  storm_hip_stormbin_read(const char* path, uint32_t max_rows, uint32_t stride,
                             uint64_t** offs, uint32_t** pos, uint32_t* n_rows,
                             uint32_t* universe);
int  storm_hip_corpus_from_stormbin(storm_hip_ctx* ctx, const char* path, uint32_t max_rows,
                                    uint32_t stride, uint32_t flags, storm_hip_corpus** out);

/* Arenas of a general corpus back to the host (layouts as storm_hip_arena;
 * EWAH marker tables: per marker its first decoded word and its offset in the
 * row's EWAH buffer; BM_ROW: bitmap slot of each row or 0xFFFFFFFF; BITMAP:
 * one row of `ld` words per slot, ld a multiple of 64; RANK_ROW / OCC_ROW:
 * index slot of each row or 0xFFFFFFFF; RANK: 2*(ceil(n_words/8)+1) words per
 * slot; OCC: ceil(ceil(n_words/8)/64) words per slot). Sizes are in bytes; a
 * download must ask for exactly the arena's size. */
enum {
    STORM_HIP_ARENA_LIST_OFF = 0, STORM_HIP_ARENA_LIST, STORM_HIP_ARENA_RUN_OFF,
    STORM_HIP_ARENA_RUN_START, STORM_HIP_ARENA_RUN_END, STORM_HIP_ARENA_EWAH_OFF,
    STORM_HIP_ARENA_EWAH, STORM_HIP_ARENA_MK_OFF, STORM_HIP_ARENA_MK_WORD,
    STORM_HIP_ARENA_MK_POS, STORM_HIP_ARENA_COMP_OFF, STORM_HIP_ARENA_COMP,
    STORM_HIP_ARENA_META, STORM_HIP_ARENA_BM_ROW, STORM_HIP_ARENA_BITMAP,
    STORM_HIP_ARENA_RANK_ROW, STORM_HIP_ARENA_RANK, STORM_HIP_ARENA_OCC_ROW,
    STORM_HIP_ARENA_OCC,
    STORM_HIP_ARENA_COUNT
};
int  storm_hip_corpus_arena_sizes(storm_hip_ctx* ctx, const storm_hip_corpus* c,
                                  uint64_t bytes[STORM_HIP_ARENA_COUNT]);
int  storm_hip_corpus_download_arena(storm_hip_ctx* ctx, const storm_hip_corpus* c,
                                     uint32_t which, void* dst, uint64_t bytes);
/* Materialise a bitmap on the device for every row that lacks one (needed by
 * the all-bitmap policy on sparse corpora). */
int  storm_hip_corpus_materialize_bitmaps(storm_hip_ctx* ctx, storm_hip_corpus* c);
void storm_hip_corpus_free(storm_hip_corpus* c);
uint32_t storm_hip_corpus_rows(const storm_hip_corpus* c);
/* Device bytes held by the corpus (arenas + result buffer). */
uint64_t storm_hip_corpus_device_bytes(const storm_hip_corpus* c);

/* ---- the all-pairs call --------------------------------------------------
 *
 * Computes |X_i n X_j| for every i < j exactly. Replaces
 *   STORM_contig_pairw_intersect_cardinality[_blocked][_list]
 *     (storm.cpp:1232-1285)                      -> policy ALL_BITMAP, SUM
 *   storm::allpairs_sum(rows, model, policy, tile)
 *     (kernels/storm_allpairs.cpp:486-675)       -> SUM, stats
 *   storm::allpairs_tiles(...)  (:677-732)       -> UPPER_TRI, original order
 */
enum {
    STORM_HIP_POLICY_ALL_BITMAP = 0,  /* every pair B x B, no span test      */
    STORM_HIP_POLICY_ANALYTIC   = 1,  /* per-tile storm_cost selection       */
    STORM_HIP_POLICY_LEARNED    = 2,  /* per-tile MLP selection (fallback:
                                         analytic when features are out of
                                         the training range)                 */
    STORM_HIP_POLICY_FIXED      = 3,  /* one cell for every eligible pair     */
    STORM_HIP_POLICY_EXHAUSTIVE = 4,  /* time every eligible cell per tile    */
    STORM_HIP_POLICY_PROBE      = 5,  /* Policy::Probe: measure, then commit
                                         (below)                             */
    STORM_HIP_POLICY_REFINE     = 6,  /* Policy::Refine: per-pair choice
                                         between a tile's two best cells     */
    STORM_HIP_POLICY_PER_PAIR   = 7   /* Policy::PerPair / Oracle: every pair
                                         selects its own cell (below)        */
};
/* PROBE (kernels/storm_allpairs.cpp:445-482 measures a tile's two best
 * candidates on 3 of its pairs and commits the faster for the rest). Timing 3
 * pairs means nothing on a GPU, whose cost is a tile's throughput, so the
 * device measures at tile granularity and commits per class of tiles: in a
 * first pass every tile with one candidate runs, and on a sample of the
 * others (at least 4 per class, then 1 in 32) the best candidate runs for
 * real and the runner-up beside it on a shadow copy that writes nothing, each
 * work item timed; tiles are classed by (best, runner-up, predicted-ratio
 * bucket); the second pass runs every pending tile with the candidate its
 * class measured faster (in CU-time). Both passes are one call (capturable).
 * REFINE (:563-588, :636-641): a tile whose rows spread by max/mean
 * cardinality >= 8 on either side and whose runner-up is predicted within 3x
 * lets each pair run the cheaper of the two as priced on the pair's own rows
 * (one wavefront per pair); other tiles run the best candidate.
 * PER_PAIR (:645-660): select_pairing on each pair's own rows over every cell
 * the two rows can run (the reference's PerPair; its Oracle makes the same
 * model choice off the clock), one wavefront per pair after the span test. */
enum {
    STORM_HIP_OUT_SUM       = 0,  /* u64 sum only                            */
    STORM_HIP_OUT_UPPER_TRI = 1,  /* u32 per pair, original row order,
                                     idx(i,j) = i*N - i(i+1)/2 + (j-i-1)     */
    STORM_HIP_OUT_UPPER_TRI_DENSITY = 2
                                  /* u32 per pair, the same packing over the
                                     positions a < b of the density order
                                     (storm_hip_corpus_order: position a holds
                                     row perm[a]) -- the coordinates in which
                                     storm::allpairs_tiles reports counts under
                                     PerTile (kernels/storm_allpairs.cpp:
                                     689-714). The engine works in this order,
                                     so every block of counts lands in whole
                                     row segments (original order scatters
                                     them). Dense corpora: identity order, the
                                     same as UPPER_TRI.                      */
};
enum {
    STORM_HIP_BB_AUTO = 0,       /* fastest exact form for the universe       */
    STORM_HIP_BB_VALU = 1,       /* AND + v_bcnt popcount                     */
    STORM_HIP_BB_MFMA = 2,       /* 0/1 int8 MFMA with i32 accumulation       */
    STORM_HIP_BB_MFMA_FP4 = 3    /* 0/1 e2m1 MFMA with f32 accumulation
                                    (exact while counts <= 2^24; universes
                                    up to 2^24 bits)                          */
};
/* Cell classes, in storm::Pairing order (kernels/storm_cost.h:47-49) plus the
 * complement cell (kernels/cell_comp.cpp). */
enum {
    STORM_HIP_CELL_BB = 0, STORM_HIP_CELL_BS, STORM_HIP_CELL_BR, STORM_HIP_CELL_BW,
    STORM_HIP_CELL_SS, STORM_HIP_CELL_SR, STORM_HIP_CELL_SW, STORM_HIP_CELL_RR,
    STORM_HIP_CELL_RW, STORM_HIP_CELL_WW, STORM_HIP_CELL_EMPTY, STORM_HIP_CELL_CB,
    STORM_HIP_NCELL
};

typedef struct {
    uint32_t policy;      /* STORM_HIP_POLICY_*                               */
    uint32_t out_mode;    /* STORM_HIP_OUT_*                                  */
    uint32_t bb_kernel;   /* STORM_HIP_BB_*                                   */
    uint32_t tile;        /* rows per selection tile (0 -> 64)               */
    uint32_t fixed_cell;  /* STORM_HIP_CELL_* for POLICY_FIXED                */
    uint32_t use_graph;   /* capture select->dispatch->reduce in a hipGraph   */
    uint32_t reserved[2];
} storm_hip_options;

typedef struct {
    uint64_t sum;                       /* sum over pairs of |Xi n Xj|        */
    uint64_t pairs;                     /* N(N-1)/2                           */
    uint64_t skipped;                   /* settled by the span test           */
    uint64_t decisions;                 /* selection decisions (tiles)        */
    uint64_t cell_pairs[STORM_HIP_NCELL];
    uint64_t cell_tiles[STORM_HIP_NCELL];
    uint64_t fallback_tiles;            /* learned -> analytic fallbacks      */
    double   ms_total;                  /* host wall time of the call         */
    double   ms_device;                 /* HIP-event time of the whole call   */
    double   ms_bb;                     /* HIP-event time of B x B launches   */
    double   ms_select;                 /* HIP-event time of selection        */
    uint64_t bb_launches;
    uint64_t probe_tiles;               /* PROBE: tiles run with both candidates */
    uint64_t probe_switched;            /* PROBE: pending tiles committed to the
                                           runner-up                          */
} storm_hip_stats;

int storm_hip_allpairs(storm_hip_ctx* ctx, storm_hip_corpus* c,
                       const storm_hip_options* opt, storm_hip_stats* stats);
/* Copy `count` packed upper-triangle counts starting at `offset` from the
 * last UPPER_TRI or UPPER_TRI_DENSITY call on this corpus (in the order that
 * call wrote). */
int storm_hip_result_copy_tri(storm_hip_ctx* ctx, const storm_hip_corpus* c,
                              uint64_t offset, uint64_t count, uint32_t* out);
/* Row sums of the last UPPER_TRI or UPPER_TRI_DENSITY result, computed on
 * the device, by original row either way: row_sum[i] = sum_{j != i}
 * |Xi n Xj|. */
int storm_hip_result_row_sums(storm_hip_ctx* ctx, const storm_hip_corpus* c,
                              uint64_t* row_sum);

/* Thresholded mode by prefix-filtered candidate generation
 * (bench/bench_threshold.cpp:210-290; Bayardo, Ma & Srikant, WWW 2007): the
 * same contract and the same reported pairs as storm_hip_allpairs_threshold
 * (exact counts of every pair with |A n B| > 0 and Jaccard >= t, read back
 * with storm_hip_result_copy_pairs), but the pairs examined are only those
 * sharing an element of their prefixes (p(X) = |X| - ceil(t|X|) + 1 elements,
 * rarest document frequency first), generated from a postings index of the
 * prefixes on the device and verified exactly. mode:
 *   STORM_HIP_PREFIX_GATED   the reference's gate (bench_threshold.cpp:170-200):
 *                            prefix filtering when its estimated work (2 sum p
 *                            + sum_e L_e(L_e-1)/2 x mean |X|) is below 0.04 of
 *                            the exact scan's (pairs x mean |X|) -- the cut
 *                            measured for the device's two paths (the
 *                            reference's CPU cut is 5) -- else, or when the
 *                            candidates outgrow one sort (2^31), the exact
 *                            thresholded engine
 *   STORM_HIP_PREFIX_ALWAYS  prefix filtering (the rows must hold lists)
 *   STORM_HIP_PREFIX_EXACT   the exact thresholded engine
 * exact_opt: options of the exact engine when it runs (NULL: analytic, 256-row
 * tiles). info (optional) reports the gate's estimate and the candidate counts;
 * under prefix filtering stats->sum is the sum of the reported pairs' counts. */
enum { STORM_HIP_PREFIX_GATED = 0, STORM_HIP_PREFIX_ALWAYS = 1, STORM_HIP_PREFIX_EXACT = 2 };
typedef struct storm_hip_prefix_info {
    double   est_ratio;          /* estimated prefix work / exact-scan work      */
    uint64_t prefix_elements;    /* sum over rows of p(X)                        */
    uint64_t candidates;         /* (row, earlier row) emissions, duplicates in  */
    uint64_t unique_candidates;  /* distinct pairs verified                      */
    uint32_t used_prefix;        /* 1: prefix filtering ran; 0: the exact engine */
} storm_hip_prefix_info;
int storm_hip_allpairs_threshold_prefix(storm_hip_ctx* ctx, storm_hip_corpus* c, double t,
                                        uint32_t mode, const storm_hip_options* exact_opt,
                                        storm_hip_stats* stats, uint64_t* n_pairs,
                                        storm_hip_prefix_info* info);

/* ---- streamed tile rows and cross blocks (general corpora) ---------------
 *
 * storm::allpairs_tiles (kernels/storm_allpairs.cpp:677-732) visits the upper
 * triangle one tile row at a time, in density order under PerTile and in row
 * order otherwise, so a consumer never holds N^2 counts. A band call computes
 * one tile row of that walk and nothing else: only the tile pairs (band, bj),
 * bj >= band, are selected and dispatched; the device holds one tile x N
 * count buffer, not the triangle.
 *   order   STORM_HIP_ORDER_DENSITY: positions of the density order (stable by
 *           cardinality, kernels/storm_allpairs.cpp:314-323; the permutation
 *           is storm_hip_corpus_order); STORM_HIP_ORDER_IDENTITY: row order.
 *   out     (rows in the band) x N u32, out[(i - band*tile) * N + j] = the
 *           count of positions i < j; entries with j <= i are 0.
 * opt->tile (64, 128, 256) is the band height. Policies: ANALYTIC, LEARNED,
 * FIXED, ALL_BITMAP (every tile B x B over materialised bitmaps). */
enum { STORM_HIP_ORDER_DENSITY = 0, STORM_HIP_ORDER_IDENTITY = 1 };
int storm_hip_allpairs_band(storm_hip_ctx* ctx, storm_hip_corpus* c,
                            const storm_hip_options* opt, uint32_t order, uint32_t band,
                            uint32_t* out, storm_hip_stats* stats);
/* Density order of a general corpus: perm[k] = the row at position k
 * (identity for a dense corpus). */
int storm_hip_corpus_order(storm_hip_ctx* ctx, const storm_hip_corpus* c, uint32_t* perm);
/* The n1 x n2 block of a corpus whose rows [0, split) are one set and
 * [split, N) another (STORM_intersect_cardinality_square, storm.cpp:1027-1046:
 * every pair across the two, none within either): only tile pairs across
 * the split are selected and dispatched. split is a multiple of opt->tile
 * (pad the first set with empty rows). stats->sum is the sum over the block;
 * out (optional, split x (N - split) u32) receives
 * out[i * (N - split) + (j - split)] = |X_i n X_j|. */
int storm_hip_allpairs_cross(storm_hip_ctx* ctx, storm_hip_corpus* c,
                             const storm_hip_options* opt, uint32_t split, uint32_t* out,
                             storm_hip_stats* stats);

/* ---- allocation-free dense call ------------------------------------------
 * No counterpart in the reference. For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc: the dense all-pairs call (as storm_hip_allpairs on
 * a STORM_contiguous_t corpus, matrix-core B x B) over rows the caller holds
 * in device memory, without any allocation by the library. d_words: n_rows
 * rows of n_words u64, row stride ld words (ld >= n_words, a multiple of 8,
 * and at least 16; the words of a row from n_words up to the next multiple of
 * 8 are read and must be zero); no row padding is needed. Outputs in device
 * memory:
 * *d_sum = sum over pairs of |Xi n Xj|; d_tri (optional) = the packed upper
 * triangle, idx(i,j) = i*N - i(i+1)/2 + (j-i-1). bb_kernel: AUTO, MFMA or
 * MFMA_FP4. The call is enqueued on the context's stream and returns
 * without waiting (storm_hip_sync waits). Limit: at most 2^20 rows (the
 * library's fixed scratch array, 4 MiB per GPU, holds the B x B row map;
 * larger calls are STORM_HIP_ERR_RANGE). Every call on one GPU shares that
 * array: calls that overlap on one GPU (two contexts or threads) race. */
int storm_hip_allpairs_contig_noalloc(storm_hip_ctx* ctx, const uint64_t* d_words,
                                      uint32_t n_rows, uint32_t n_words, uint32_t ld,
                                      uint32_t bb_kernel, uint32_t* d_tri,
                                      unsigned long long* d_sum);

/* ---- rectangular form -----------------------------------------------------
 *
 * |Xi n Yj| for every row i of x and every row j of y: the two-input form the
 * reference plans as the STORM_xy family (RESEARCH_PLAN.md:448-449), of which
 * the upper triangle is the special case x == y. Two dense corpora of the
 * same universe; matrix-core B x B only (bb_kernel AUTO, MFMA or MFMA_FP4,
 * else STORM_HIP_ERR_UNSUPPORTED). *sum receives the total over all n_x * n_y
 * pairs; out (host, n_x * n_y u32, row-major, out[i * n_y + j]) may be NULL. */
int storm_hip_allpairs_xy(storm_hip_ctx* ctx, const storm_hip_corpus* x,
                          const storm_hip_corpus* y, uint32_t bb_kernel, uint64_t* sum,
                          uint32_t* out);

/* ---- thresholded mode -----------------------------------------------------
 *
 * Exact |Xi n Xj| for exactly the pairs i < j whose Jaccard similarity reaches
 * t, the opt-in contract of the reference's thresholded mode
 * (bench/bench_threshold.cpp, README.md:114-161): a pair is reported when
 * |Xi n Xj| > 0 and |Xi n Xj| / (|Xi| + |Xj| - |Xi n Xj|) >= t, evaluated in
 * double precision as there. The reference generates candidates with a
 * prefix-filter index; here the selection kernel drops every tile pair whose
 * cardinality ranges bound the Jaccard below t (min/max length filter, exact
 * in the same double arithmetic), the cells skip pairs failing the same bound,
 * and surviving pairs are counted exactly and appended on the device.
 * General corpora and the ANALYTIC / LEARNED / FIXED policies (else
 * STORM_HIP_ERR_UNSUPPORTED); t outside (0, 1] is STORM_HIP_ERR_INVALID.
 * *n_pairs receives the number of reported pairs; they are kept on the corpus
 * sorted by (i, j) in original row ids until the next thresholded call.
 * stats->sum is the sum over the pairs the call counted (not every pair). */
typedef struct {
    uint32_t i, j;      /* original row ids, i < j */
    uint32_t count;     /* |Xi n Xj| */
} storm_hip_pair;
int storm_hip_allpairs_threshold(storm_hip_ctx* ctx, storm_hip_corpus* c,
                                 const storm_hip_options* opt, double t,
                                 storm_hip_stats* stats, uint64_t* n_pairs);
int storm_hip_result_copy_pairs(storm_hip_ctx* ctx, const storm_hip_corpus* c,
                                uint64_t offset, uint64_t count, storm_hip_pair* out);

/* ---- launch shapes and the autotuner -------------------------------------
 * No reference counterpart (the reference's kernels have fixed shapes).
 * The matrix-core B x B kernel's launch shape: packed words per LDS stage
 * (4, 8 or 16; 16 with 8 waves only, 144 KB of LDS), tile order (0 = row-major upper triangle, g > 0 = bands of g tile
 * rows walked column by column, each XCD taking a contiguous range) and the
 * K-split target (blocks per CU when a corpus has fewer 256 x 256 tiles than
 * that) and the waves per workgroup (8: two per SIMD, 128 x 64 pairs each;
 * 4: one per SIMD, 128 x 128 pairs each). A context starts from {16, 4, 2, 8}. */
typedef struct {
    uint32_t stage_words;
    uint32_t band;
    uint32_t blocks_per_cu;
    uint32_t waves;
} storm_hip_bb_config;
int storm_hip_set_bb_config(storm_hip_ctx* ctx, const storm_hip_bb_config* cfg);
int storm_hip_get_bb_config(storm_hip_ctx* ctx, storm_hip_bb_config* cfg);

typedef struct {
    uint32_t tile;                 /* best selection tile (general corpora; 0 for dense) */
    uint32_t n_tried;              /* candidates timed                                    */
    storm_hip_bb_config bb;        /* best B x B launch shape (now the context's)         */
    double ms_best;                /* device time of one call with the choice             */
    double ms_default;             /* device time with the starting shape and tile        */
} storm_hip_tune_result;
/* Time candidate shapes on this corpus with `opt`'s policy and B x B kernel
 * (each candidate runs the whole call once, sums only, after one warm-up)
 * and keep the fastest: general corpora try selection tiles 64, 128 and 256,
 * then, when the best tile puts B x B on the matrix cores (tile 256) or the
 * corpus is dense, every stage width x tile order x K-split target. The
 * context's B x B shape becomes the best one; pass result.tile as the tile
 * of later calls. Counts never depend on the shape. */
int storm_hip_autotune(storm_hip_ctx* ctx, storm_hip_corpus* c, const storm_hip_options* opt,
                       storm_hip_tune_result* result);

/* ---- batches of independent corpora --------------------------------------
 * One call over many general corpora: every corpus is selected per tile
 * (analytic, learned or fixed policy), the tiles of all corpora share one
 * device queue per cell class, and each class runs as ONE persistent launch
 * for the whole batch (with use_graph, the whole sequence is one hipGraph).
 * The reference runs independent problems one call at a time
 * (kernels/storm_allpairs.h:196-214, storm.h:218); this has no single
 * counterpart there. `sums[k]` (optional) receives corpus k's sum;
 * `latency_ms[k]` (optional) the time from the start of the call's launches
 * to the end of corpus k's last tile (0 when it has no tile). With
 * OUT_UPPER_TRI every corpus keeps its own triangle (result_copy_tri). The
 * corpora must stay alive while the batch exists. */
typedef struct storm_hip_batch storm_hip_batch;
int storm_hip_batch_create(storm_hip_ctx* ctx, storm_hip_corpus* const* corpora, uint32_t n,
                           storm_hip_batch** out);
int storm_hip_batch_allpairs(storm_hip_ctx* ctx, storm_hip_batch* b,
                             const storm_hip_options* opt, storm_hip_stats* stats,
                             uint64_t* sums, double* latency_ms);
void storm_hip_batch_free(storm_hip_batch* b);

/* ---- analytic cost model on the device -----------------------------------
 * CostModel rates (kernels/storm_cost.h:60-92) used by POLICY_ANALYTIC, in
 * ns per work unit, indexed by STORM_HIP_CELL_* (BB..WW); a rate <= 0 removes
 * the cell from the candidate set, as select_pairing does (storm_cost.cpp:456). */
typedef struct {
    double ns_per_unit[10];
    double ns_fixed;
    double ns_per_occ_word;
    double ns_probe_hot, ns_probe_cold;
    double bb_zonemap;   /* 1: reference pricing of a zone-mapped B x B; 0: Theta(m) */
} storm_hip_cost_model;
int storm_hip_set_cost_model(storm_hip_ctx* ctx, const storm_hip_cost_model* m);
int storm_hip_get_cost_model(storm_hip_ctx* ctx, storm_hip_cost_model* m);
/* calibrate (kernels/storm_cost.cpp:556-769) on the device, on this corpus's
 * own tiles (the reference's calibrate_on_rows idea, :779-895): time every
 * eligible cell on up to max_tiles sampled tile pairs and set each cell's rate
 * to the median over tiles of time / (pairs x work_units); the device B x B is
 * priced per word (bb_zonemap = 0); ns_fixed and the occ/probe terms are 0.
 * The context's model is replaced; `out` (optional) receives it. */
int storm_hip_calibrate(storm_hip_ctx* ctx, storm_hip_corpus* c, uint32_t tile,
                        uint32_t max_tiles, uint64_t seed, uint32_t repeats,
                        storm_hip_cost_model* out);
/* The same fit from timings already drawn by storm_hip_selector_sample on
 * this corpus at this tile width (n tiles: log2_ns, mask, tile_index as that
 * call returned them), so one sampling pass serves both models. */
int storm_hip_calibrate_samples(storm_hip_ctx* ctx, storm_hip_corpus* c, uint32_t tile,
                                uint32_t n, const float* log2_ns, const uint8_t* mask,
                                const uint32_t* tile_index, storm_hip_cost_model* out);
/* The same fit pooled over samples of several corpora (e.g. the instances of
 * a batch): sample k was drawn by storm_hip_selector_sample on
 * corpora[corpus_index[k]] at this tile width; each cell's rate is the median
 * over all its samples. */
int storm_hip_calibrate_corpora(storm_hip_ctx* ctx, storm_hip_corpus* const* corpora,
                                uint32_t n_corpora, uint32_t tile, uint32_t n,
                                const float* log2_ns, const uint8_t* mask,
                                const uint32_t* tile_index, const uint32_t* corpus_index,
                                storm_hip_cost_model* out);
/* Per-tile decisions of the last selecting call: `cells[t]` for tile pair t
 * in row-major upper-triangle tile order. */
int storm_hip_last_decisions(storm_hip_ctx* ctx, const storm_hip_corpus* c,
                             uint8_t* cells, uint64_t n_tiles);

/* ---- learned per-tile selector -------------------------------------------
 * No reference counterpart: the reference selects with the analytic model
 * only (kernels/storm_cost.cpp:422-465) and measures candidates at run time in
 * its Probe policy (kernels/storm_allpairs.cpp:445-482). Here a small MLP
 * (STORM_HIP_MLP_FEATURES tile features -> 32 -> 32 -> 10 cells, ReLU, fp32)
 * predicts log2 of each cell's time on a tile and POLICY_LEARNED takes the
 * cheapest eligible cell; tiles whose features leave the training range use
 * the analytic model. Parameters are packed W1[32][F] b1 W2[32][32] b2
 * W3[10][32] b3; `norm` is mean[F], inv_std[F], lo[F], hi[F]. */
#define STORM_HIP_MLP_FEATURES 16
int storm_hip_mlp_shape(int* features, int* h1, int* h2, int* outputs, int* n_params);
/* Time every eligible cell on up to max_tiles sampled non-empty tile pairs
 * (tile rows each), `repeats` times (min kept). Each timing launch runs the
 * tiles' work items as production queues them, filled to every workgroup slot
 * with items of other tiles (run, not timed), and a tile's cost is its
 * items' CU-time (latency over the workgroups sharing a CU); per-pair cells
 * are timed on one 64 x 64 block and scaled by the pair count. Outputs per
 * sampled tile: features[F], log2 of that cost in ns per cell (mask=1 where
 * the cell was eligible and timed), the tile-pair index and its pair count. */
int storm_hip_selector_sample(storm_hip_ctx* ctx, storm_hip_corpus* c, uint32_t tile,
                              uint32_t max_tiles, uint64_t seed, uint32_t repeats,
                              float* features, float* log2_ns, uint8_t* mask,
                              uint32_t* tile_index, uint32_t* n_sampled);
/* Exploration (the production-mix alternative to storm_hip_selector_sample):
 * one all-pairs call, exact as any other, in which a seeded share `fraction`
 * of the non-empty tile pairs runs a cell drawn from the cells the learned
 * selector may rank on it and the others run the analytic choice, and every
 * work item's time is recorded as the launches run, alongside the other cell
 * classes on their streams. Outputs, per tile pair that ran a
 * cell (at most `capacity`): its features, log2 of its cost in ns -- the
 * CU-time of its items, wall time over its class's workgroups per CU -- for
 * the cell it ran (mask 1 there only), and its tile-pair index; the format
 * storm_hip_selector_train and storm_hip_calibrate_samples take; *sum
 * (optional) the call's exact total. Repeated calls with other seeds label
 * other cells of the same tiles. */
int storm_hip_selector_explore(storm_hip_ctx* ctx, storm_hip_corpus* c, uint32_t tile,
                               uint64_t seed, double fraction, uint32_t capacity, float* features,
                               float* log2_ns, uint8_t* mask, uint32_t* tile_index,
                               uint32_t* n_sampled, uint64_t* sum);
/* The same over a batch (corpus_index[k]: the batch member of sample k;
 * sums, optional: each member's total). */
int storm_hip_batch_explore(storm_hip_ctx* ctx, storm_hip_batch* b, uint32_t tile, uint64_t seed,
                            double fraction, uint32_t capacity, float* features, float* log2_ns,
                            uint8_t* mask,
                            uint32_t* tile_index, uint32_t* corpus_index, uint32_t* n_sampled,
                            uint64_t* sums);
/* Train on the device (He init from `seed`, Adam, minibatches of `batch`);
 * the trained model then serves POLICY_LEARNED on this context. */
int storm_hip_selector_train(storm_hip_ctx* ctx, const float* features, const float* log2_ns,
                             const uint8_t* mask, uint32_t n, uint32_t epochs, uint32_t batch,
                             float lr, uint64_t seed, float* final_loss);
/* Network outputs (log2 ns per cell) for n raw feature rows, normalised with
 * the model's stored statistics. */
int storm_hip_selector_predict(storm_hip_ctx* ctx, const float* features, uint32_t n, float* out);
int storm_hip_selector_set_params(storm_hip_ctx* ctx, const float* params, const float* norm);
int storm_hip_selector_get_params(storm_hip_ctx* ctx, float* params, float* norm);
/* Gradient of the masked mean-square loss at the current parameters over
 * (already normalised) inputs; and one Adam step (b1 .9, b2 .999, eps 1e-8)
 * with the last gradient. Exposed so the device math can be checked. */
int storm_hip_selector_grad(storm_hip_ctx* ctx, const float* x, const float* t,
                            const uint8_t* mask, uint32_t n, float* grad, float* loss);
int storm_hip_selector_adam_step(storm_hip_ctx* ctx, float lr);

#ifdef __cplusplus
}
#endif
#endif
