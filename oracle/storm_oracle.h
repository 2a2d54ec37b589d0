/*
 * CPU oracle for exact all-pairs set-intersection cardinality.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference
 * algorithm, used by tests/, by the smoke check and by the benchmark's CPU
 * baseline leg as the CHECKER. It is never linked into, loaded by or called
 * from the product library; the product fails loudly without its HIP build.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the StormBitmaps source tree).
 */
#ifndef STORM_ORACLE_H_
#define STORM_ORACLE_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded stream: splitmix64 (kernels/storm_gen.h:84-102) ------------- */
uint64_t orc_rng_next(uint64_t* s);
uint32_t orc_rng_below(uint64_t* s, uint32_t n);
double   orc_rng_unit(uint64_t* s);
uint64_t orc_rng_seed(uint64_t seed);        /* seed==0 -> golden constant */

/* Row-major Bernoulli(p) bitmap matrix from ONE stream Rng(seed): bit b of row
 * r is set iff the (r*n_bits+b)-th unit() draw is < p. n_words = ceil(bits/64),
 * bits past n_bits in the last word stay zero. `row0` skips the first row0 rows
 * of the stream so a sample of rows can be regenerated alone (the stream is
 * counter-based, so skipping is O(1)). */
void orc_gen_bernoulli_rows(uint64_t seed, double p, uint32_t n_bits,
                            uint32_t row0, uint32_t n_rows, uint64_t* out_words);

/* ---- popcount and the per-pair oracle (kernels/oracle.cpp:21-38) -------- */
uint32_t orc_popcount_bitloop(uint64_t x);
/* |A n B| from two sorted position lists, restricted to the first
 * min(nw_a, nw_b) words, by rebuilding both bitmaps and bit-loop counting. */
uint64_t orc_intersect_lists(const uint32_t* a, uint32_t na, uint32_t nw_a,
                             const uint32_t* b, uint32_t nb, uint32_t nw_b);
/* Same quantity by a two-pointer merge; cross-checked against the above. */
uint64_t orc_intersect_merge(const uint32_t* a, uint32_t na,
                             const uint32_t* b, uint32_t nb);
/* Same quantity from two dense bitmaps of nw words (bit-loop popcount). */
uint64_t orc_intersect_words(const uint64_t* a, const uint64_t* b, uint32_t nw);

/* ---- representations (kernels/storm_repr.{h,cpp}) ----------------------- */
typedef struct {
    uint32_t cardinality, n_runs, n_words, n_nonzero_w, first_set, last_set;
    uint8_t  has_rank, has_occ;
} orc_meta;                                  /* storm_repr.h:213-231 */

typedef struct {
    uint64_t* bitmap;   uint32_t n_words;
    uint64_t* rank;     uint32_t n_rank;     /* 2*(nblk+1) words, or 0 */
    uint64_t* occ;      uint32_t n_occ;      /* ceil(bins/64) words, or 0 */
    uint32_t* list;     uint32_t n_list;
    uint32_t* run_start; uint32_t* run_end; uint32_t n_runs;
    uint64_t* ewah;     uint32_t n_ewah;
    uint32_t* comp_list; uint32_t n_comp;
    uint32_t* comp_start; uint32_t* comp_end; uint32_t n_comp_runs;
    orc_meta  meta;
} orc_row;

/* build_row (storm_repr.cpp:102-207), including the index gates: rank only if
 * mean run length >= 512, zone map only if n*512 > universe, complement only
 * if 2n > universe. Positions must be sorted and distinct. Returns 0 or -1. */
int  orc_build_row(orc_row* out, const uint32_t* pos, uint32_t n, uint32_t universe);
void orc_free_row(orc_row* r);
/* build_rank (storm_repr.cpp:25-45) unconditionally, rank9 layout. */
void orc_build_rank(const uint64_t* w, uint32_t nw, uint64_t* out /* 2*(nblk+1) */);
uint64_t orc_rank_at(const uint64_t* w, uint32_t nw, const uint64_t* rank, uint32_t x);
/* build_occ (storm_repr.cpp:47-59): 1 bit per `bin` words. */
void orc_build_occ(const uint64_t* w, uint32_t nw, uint32_t bin, uint64_t* out);
/* build_ewah (storm_repr.cpp:78-100); returns words written (<= 2*nw+1). */
uint32_t orc_build_ewah(const uint64_t* w, uint32_t nw, uint64_t* out);
/* ewah_decode (storm_repr.cpp:61-74); returns decoded words. */
uint32_t orc_ewah_decode(const uint64_t* buf, uint32_t n, uint64_t* out);

/* ---- cost model (kernels/storm_cost.{h,cpp}) ---------------------------- */
enum { ORC_BB, ORC_BS, ORC_BR, ORC_BW, ORC_SS, ORC_SR, ORC_SW, ORC_RR, ORC_RW,
       ORC_WW, ORC_EMPTY, ORC_NCELL };      /* storm_cost.h:47-49 */
typedef struct {
    double ns_per_unit[ORC_NCELL];
    double ns_fixed, ns_per_occ_word, ns_probe_hot, ns_probe_cold, l2_bytes;
} orc_cost_model;                            /* storm_cost.h:60-92 */
void   orc_default_model(orc_cost_model* m); /* storm_cost.cpp:482-496 */
double orc_work_units(int p, const orc_meta* a, const orc_meta* b, double br_margin);
double orc_predict(const orc_cost_model* m, int p, const orc_meta* a, const orc_meta* b,
                   double probe_scale, double br_margin, double rr_scale);
int    orc_select_pairing(const orc_cost_model* m, const orc_meta* a, const orc_meta* b,
                          double* predicted);  /* default knobs: 0.05, 16, 1 */

/* ---- driver helpers (kernels/storm_allpairs.cpp) ------------------------ */
/* tile_meta (:281-312): mean card/runs/nnz (integer division), min first,
 * max last (first=0 if all empty), AND of has_rank/has_occ, max n_words. */
void orc_tile_meta(const orc_meta* metas, const uint32_t* order, uint32_t lo, uint32_t hi,
                   orc_meta* out);
/* density_order (:316-323): stable sort of row ids by cardinality. */
void orc_density_order(const orc_meta* metas, uint32_t n, uint32_t* order);

/* ---- whole-matrix references -------------------------------------------- */
/* Upper-triangle counts in ORIGINAL row order, idx(i,j)=i*N-i(i+1)/2+(j-i-1),
 * from CSR lists (offs has n+1 entries). Returns the sum. Multi-threaded over
 * rows with `threads` workers (0 = 1). */
uint64_t orc_allpairs_lists(const uint32_t* offs, const uint32_t* pos, uint32_t n,
                            uint32_t* out_tri /* may be NULL */, int threads);
/* Same from a dense row-major bitmap matrix (n rows x nw words). */
uint64_t orc_allpairs_words(const uint64_t* words, uint32_t n, uint32_t nw,
                            uint32_t* out_tri /* may be NULL */, int threads);
/* Closed forms from column counts c_b (the number of rows holding bit b):
 * total = sum_b C(c_b,2); row_sum[i] = sum_{b in X_i} (c_b - 1). */
void     orc_column_counts_words(const uint64_t* words, uint32_t n, uint32_t nw,
                                 uint32_t* col /* nw*64 */);
uint64_t orc_total_from_columns(const uint32_t* col, uint64_t n_cols);
/* Column counts and (optional) row sums as above, on `threads` threads. */
void     orc_column_closed_forms_words(const uint64_t* words, uint32_t n, uint32_t nw,
                                       uint32_t* col, uint64_t* row_sum, int threads);
/* The triangle's rows [r0, r1): every pair (i, j), r0 <= i < r1, i < j < n,
 * written in triangle order from out[0] (out may be NULL); returns their sum. */
uint64_t orc_band_words(const uint64_t* words, uint32_t n, uint32_t nw, uint32_t r0,
                        uint32_t r1, uint32_t* out, int threads);
void     orc_row_sums_words(const uint64_t* words, uint32_t n, uint32_t nw,
                            const uint32_t* col, uint64_t* row_sum);
/* Row sums implied by a packed upper-triangle output: row_sum[i] =
 * sum_{j != i} out(min(i,j), max(i,j)). */
void     orc_row_sums_from_tri(const uint32_t* tri, uint32_t n, uint64_t* row_sum);

/* ---- learned-selector math: float64 MLP reference ----------------------- */
/* Two hidden layers with ReLU, linear output; row-major weights
 * W1[h1][f], W2[h2][h1], W3[o][h2]. Loss = mean over batch and outputs of
 * (y - t)^2 / 2 restricted to mask==1 targets. Gradients into g*. */
double orc_mlp_forward_backward(int f, int h1, int h2, int o, int batch,
                                const double* x, const double* t, const uint8_t* mask,
                                const double* W1, const double* b1,
                                const double* W2, const double* b2,
                                const double* W3, const double* b3,
                                double* y /* batch*o, may be NULL */,
                                double* gW1, double* gb1, double* gW2, double* gb2,
                                double* gW3, double* gb3);
/* One Adam step (Kingma & Ba 2015, Algorithm 1) on n parameters, step t>=1. */
void orc_adam_step(int n, double* p, const double* g, double* m, double* v, int t,
                   double lr, double beta1, double beta2, double eps);

/* 1 when this CPU runs AVX-512 F/BW/VL/VPOPCNTDQ (the reference build that
 * uses them may be loaded), from the CPUID bits. */
int orc_cpu_has_avx512_vpopcntdq(void);

#ifdef __cplusplus
}
#endif
#endif
