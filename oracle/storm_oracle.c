/*
 * CPU oracle for exact all-pairs set-intersection cardinality.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference algorithm,
 * used as the checker by tests/, the smoke check and the benchmark's CPU
 * baseline leg. Never linked into the product library.
 *
 * Parity pinning of each part is listed beside it; "pinned by" names the test
 * that would fail on a plausible mistake here.
 */
#define _GNU_SOURCE
#include "storm_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* splitmix64: kernels/storm_gen.h:84-102. Pinned by the known config-1 sum  */
/* (54,935,083) and by the compiled reference's Rng in tests.                */
/* ------------------------------------------------------------------------ */
#define ORC_GOLDEN 0x9E3779B97F4A7C15ull

uint64_t orc_rng_seed(uint64_t seed) { return seed ? seed : ORC_GOLDEN; }

static inline uint64_t orc_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t orc_rng_next(uint64_t* s) {
    *s += ORC_GOLDEN;
    return orc_mix(*s);
}

uint32_t orc_rng_below(uint64_t* s, uint32_t n) {
    if (n == 0) return 0;
    const uint64_t lim = UINT64_MAX - (UINT64_MAX % n);
    uint64_t r;
    do { r = orc_rng_next(s); } while (r >= lim);
    return (uint32_t)(r % n);
}

double orc_rng_unit(uint64_t* s) {
    return (double)(orc_rng_next(s) >> 11) * (1.0 / 9007199254740992.0);
}

void orc_gen_bernoulli_rows(uint64_t seed, double p, uint32_t n_bits,
                            uint32_t row0, uint32_t n_rows, uint64_t* out) {
    const uint32_t nw = (n_bits + 63u) / 64u;
    /* The k-th draw of Rng(seed) is mix(seed + (k+1)*golden). */
    uint64_t s = orc_rng_seed(seed) + (uint64_t)row0 * n_bits * ORC_GOLDEN;
    for (uint32_t r = 0; r < n_rows; ++r) {
        uint64_t* w = out + (size_t)r * nw;
        memset(w, 0, (size_t)nw * 8);
        for (uint32_t b = 0; b < n_bits; ++b)
            if (orc_rng_unit(&s) < p) w[b >> 6] |= 1ull << (b & 63);
    }
}

/* ------------------------------------------------------------------------ */
/* Oracle: kernels/oracle.cpp:21-38.                                         */
/* ------------------------------------------------------------------------ */
uint32_t orc_popcount_bitloop(uint64_t x) {
    uint32_t n = 0;
    while (x) { n += (uint32_t)(x & 1u); x >>= 1; }
    return n;
}

uint64_t orc_intersect_lists(const uint32_t* a, uint32_t na, uint32_t nw_a,
                             const uint32_t* b, uint32_t nb, uint32_t nw_b) {
    const uint32_t nw = nw_a < nw_b ? nw_a : nw_b;
    uint64_t* ba = (uint64_t*)calloc(nw ? nw : 1, 8);
    uint64_t* bb = (uint64_t*)calloc(nw ? nw : 1, 8);
    for (uint32_t i = 0; i < na; ++i) if ((a[i] >> 6) < nw) ba[a[i] >> 6] |= 1ull << (a[i] & 63);
    for (uint32_t i = 0; i < nb; ++i) if ((b[i] >> 6) < nw) bb[b[i] >> 6] |= 1ull << (b[i] & 63);
    uint64_t c = 0;
    for (uint32_t k = 0; k < nw; ++k) c += orc_popcount_bitloop(ba[k] & bb[k]);
    free(ba); free(bb);
    return c;
}

uint64_t orc_intersect_merge(const uint32_t* a, uint32_t na, const uint32_t* b, uint32_t nb) {
    uint64_t c = 0;
    uint32_t i = 0, j = 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) ++i;
        else if (b[j] < a[i]) ++j;
        else { ++c; ++i; ++j; }
    }
    return c;
}

uint64_t orc_intersect_words(const uint64_t* a, const uint64_t* b, uint32_t nw) {
    uint64_t c = 0;
    for (uint32_t k = 0; k < nw; ++k) c += orc_popcount_bitloop(a[k] & b[k]);
    return c;
}

/* ------------------------------------------------------------------------ */
/* Representations: kernels/storm_repr.cpp. Pinned by round trips (bitmap,   */
/* runs, EWAH decode, rank at every word) and by golden words from the       */
/* compiled reference.                                                       */
/* ------------------------------------------------------------------------ */
void orc_build_rank(const uint64_t* w, uint32_t nw, uint64_t* out) {
    const uint32_t nblk = (nw + 7) / 8;                 /* storm_repr.cpp:25-45 */
    uint64_t acc = 0;
    for (uint32_t b = 0; b < nblk; ++b) {
        out[2 * b] = acc;
        uint64_t sub = 0, rel = 0;
        for (uint32_t k = 0; k < 8; ++k) {
            const uint32_t wi = b * 8 + k;
            if (k >= 1) sub |= (rel & 0x1FF) << (9 * (k - 1));
            if (wi < nw) rel += orc_popcount_bitloop(w[wi]);
        }
        out[2 * b + 1] = sub;
        acc += rel;
    }
    out[2 * nblk] = acc;
    out[2 * nblk + 1] = 0;
}

uint64_t orc_rank_at(const uint64_t* w, uint32_t nw, const uint64_t* rank, uint32_t x) {
    const uint32_t wi = x >> 6;                          /* storm_repr.h:129-137 */
    if (wi >= nw) return rank[2 * ((nw + 7) / 8)];
    const uint32_t blk = wi >> 3, k = wi & 7;
    uint64_t r = rank[2 * blk];
    if (k) r += (rank[2 * blk + 1] >> (9 * (k - 1))) & 0x1FF;
    return r + orc_popcount_bitloop(w[wi] & ((1ull << (x & 63)) - 1));
}

void orc_build_occ(const uint64_t* w, uint32_t nw, uint32_t bin, uint64_t* out) {
    if (bin == 0) bin = 8;                               /* storm_repr.cpp:47-59 */
    const uint32_t bins = (nw + bin - 1) / bin;
    memset(out, 0, (size_t)((bins + 63) / 64) * 8);
    for (uint32_t b = 0; b < bins; ++b) {
        uint64_t any = 0;
        for (uint32_t k = b * bin; k < (b + 1) * bin && k < nw; ++k) any |= w[k];
        if (any) out[b >> 6] |= 1ull << (b & 63);
    }
}

uint32_t orc_build_ewah(const uint64_t* w, uint32_t nw, uint64_t* out) {
    uint32_t n = 0, i = 0;                               /* storm_repr.cpp:78-100 */
    while (i < nw) {
        int fv = (w[i] == ~0ull);
        uint64_t fl = 0;
        if (w[i] == 0 || fv) {
            const uint64_t pat = fv ? ~0ull : 0ull;
            while (i < nw && w[i] == pat && fl < 0x7FFFFFFFull) { ++i; ++fl; }
        } else {
            fv = 0;
        }
        const uint32_t lit0 = i;
        while (i < nw && w[i] != 0 && w[i] != ~0ull) ++i;
        const uint64_t ll = i - lit0;
        out[n++] = (uint64_t)fv | (fl << 1) | (ll << 32);
        for (uint64_t k = 0; k < ll; ++k) out[n++] = w[lit0 + k];
    }
    if (n == 0) out[n++] = 0;
    return n;
}

uint32_t orc_ewah_decode(const uint64_t* buf, uint32_t n, uint64_t* out) {
    uint32_t o = 0, i = 0;                               /* storm_repr.cpp:61-74 */
    while (i < n) {
        const uint64_t m = buf[i++];
        const uint64_t fl = (m >> 1) & 0x7FFFFFFFull, ll = m >> 32;
        const uint64_t fv = (m & 1) ? ~0ull : 0ull;
        for (uint64_t k = 0; k < fl; ++k) out[o++] = fv;
        for (uint64_t k = 0; k < ll; ++k) out[o++] = buf[i++];
    }
    return o;
}

void orc_free_row(orc_row* r) {
    free(r->bitmap); free(r->rank); free(r->occ); free(r->list);
    free(r->run_start); free(r->run_end); free(r->ewah);
    free(r->comp_list); free(r->comp_start); free(r->comp_end);
    memset(r, 0, sizeof *r);
}

static void orc_runs_of(const uint32_t* pos, uint32_t n, uint32_t** st, uint32_t** en,
                        uint32_t* nr) {
    *st = (uint32_t*)malloc((n ? n : 1) * 4);
    *en = (uint32_t*)malloc((n ? n : 1) * 4);
    uint32_t k = 0;
    for (uint32_t i = 0; i < n;) {
        uint32_t j = i;
        while (j + 1 < n && pos[j + 1] == pos[j] + 1) ++j;
        (*st)[k] = pos[i]; (*en)[k] = pos[j] + 1; ++k;
        i = j + 1;
    }
    *nr = k;
}

int orc_build_row(orc_row* out, const uint32_t* pos, uint32_t n, uint32_t universe) {
    memset(out, 0, sizeof *out);                         /* storm_repr.cpp:102-207 */
    const uint32_t nw = (universe + 63u) / 64u;
    for (uint32_t i = 0; i < n; ++i)
        if (pos[i] >= universe || (i && pos[i] <= pos[i - 1])) return -1;
    out->n_words = nw;
    out->bitmap = (uint64_t*)calloc(nw ? nw : 1, 8);
    out->list = (uint32_t*)malloc((n ? n : 1) * 4);
    memcpy(out->list, pos, (size_t)n * 4);
    out->n_list = n;
    for (uint32_t i = 0; i < n; ++i) out->bitmap[pos[i] >> 6] |= 1ull << (pos[i] & 63);
    orc_runs_of(pos, n, &out->run_start, &out->run_end, &out->n_runs);
    out->ewah = (uint64_t*)malloc(((size_t)2 * nw + 2) * 8);
    out->n_ewah = orc_build_ewah(out->bitmap, nw, out->ewah);
    if ((uint64_t)n * 2 > universe) {                    /* complement gate :139 */
        out->comp_list = (uint32_t*)malloc((size_t)(universe - n + 1) * 4);
        uint32_t k = 0, prev = 0;
        for (uint32_t i = 0; i < n; ++i) {
            for (uint32_t p = prev; p < pos[i]; ++p) out->comp_list[k++] = p;
            prev = pos[i] + 1;
        }
        for (uint32_t p = prev; p < universe; ++p) out->comp_list[k++] = p;
        out->n_comp = k;
        orc_runs_of(out->comp_list, k, &out->comp_start, &out->comp_end, &out->n_comp_runs);
    }
    const uint64_t mean_rl = out->n_runs ? (uint64_t)n / out->n_runs : 0;
    if (mean_rl >= 512) {                                /* rank gate :186 */
        const uint32_t nblk = (nw + 7) / 8;
        out->n_rank = 2 * (nblk + 1);
        out->rank = (uint64_t*)malloc((size_t)out->n_rank * 8);
        orc_build_rank(out->bitmap, nw, out->rank);
    }
    if ((uint64_t)n * 512 > (uint64_t)universe) {        /* occ gate :188-190 */
        const uint32_t bins = (nw + 7) / 8;
        out->n_occ = (bins + 63) / 64;
        out->occ = (uint64_t*)malloc((size_t)(out->n_occ ? out->n_occ : 1) * 8);
        orc_build_occ(out->bitmap, nw, 8, out->occ);
    }
    uint32_t nnz = 0;
    for (uint32_t k = 0; k < nw; ++k) nnz += (out->bitmap[k] != 0);
    out->meta.cardinality = n;
    out->meta.n_runs = out->n_runs;
    out->meta.n_words = nw;
    out->meta.n_nonzero_w = nnz;
    out->meta.first_set = n ? pos[0] : UINT32_MAX;
    out->meta.last_set = n ? pos[n - 1] : 0;
    out->meta.has_rank = out->rank != NULL;
    out->meta.has_occ = out->occ != NULL;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Cost model: kernels/storm_cost.cpp:253-655 (default knobs:                */
/* STORM_PROBE_SCALE=0.05, STORM_BR_ORIENT=16, STORM_RR_SCALE=1). Pinned by  */
/* golden decisions and predictions from the compiled reference.             */
/* ------------------------------------------------------------------------ */
static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }
static uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }

void orc_default_model(orc_cost_model* m) {
    memset(m, 0, sizeof *m);                             /* storm_cost.cpp:482-496 */
    m->ns_per_unit[ORC_BB] = 0.116; m->ns_per_unit[ORC_BS] = 0.30;
    m->ns_per_unit[ORC_BR] = 0.60;  m->ns_per_unit[ORC_BW] = 0.55;
    m->ns_per_unit[ORC_SS] = 0.22;  m->ns_per_unit[ORC_SR] = 0.35;
    m->ns_per_unit[ORC_SW] = 0.40;  m->ns_per_unit[ORC_RR] = 0.30;
    m->ns_per_unit[ORC_RW] = 0.45;  m->ns_per_unit[ORC_WW] = 0.50;
    m->ns_fixed = 2.0; m->ns_per_occ_word = 0.30;
}

static double orc_bb_expected_bins(const orc_meta* a, const orc_meta* b) {
    const double bins = dmax(1.0, (double)umax(a->n_words, b->n_words) / 8.0);
    const double fa = dmin(1.0, (double)a->n_nonzero_w / umax(1u, a->n_words));
    const double fb = dmin(1.0, (double)b->n_nonzero_w / umax(1u, b->n_words));
    const double pa = dmin(1.0, 8.0 * fa), pb = dmin(1.0, 8.0 * fb);
    return dmax(1.0, bins * pa * pb * 8.0);              /* storm_cost.cpp:63-75 */
}

double orc_work_units(int p, const orc_meta* a, const orc_meta* b, double margin) {
    const double sa = a->cardinality, sb = b->cardinality;   /* :195-333 */
    const double ra = a->n_runs, rb = b->n_runs;
    const double wa = 2.0 * ra + 1.0 + (double)a->n_nonzero_w;
    const double wb = 2.0 * rb + 1.0 + (double)b->n_nonzero_w;
    const int a_sparse = sa <= sb;
    const double s_card = a_sparse ? sa : sb;
    const double d_runs = a_sparse ? rb : ra;
    const double d_ewah = a_sparse ? wb : wa;
    switch (p) {
    case ORC_BB: return orc_bb_expected_bins(a, b);
    case ORC_BS: return s_card;
    case ORC_BR: {
        const double dr = a_sparse ? rb : ra, sr = a_sparse ? ra : rb;
        const int sw = margin > 0.0 && dr * margin <= sr;
        return sw ? dr : sr;
    }
    case ORC_BW: {
        const double dr = a_sparse ? rb : ra, sr = a_sparse ? ra : rb;
        const int sw = margin > 0.0 && dr * margin <= sr;
        return sw ? (2.0 * dr + 1.0 + (a_sparse ? b->n_nonzero_w : a->n_nonzero_w))
                  : (2.0 * sr + 1.0 + (a_sparse ? a->n_nonzero_w : b->n_nonzero_w));
    }
    case ORC_SS: {
        const uint32_t mx = (uint32_t)dmax(2.0, dmax(sa, sb));
        return dmin(sa, sb) * (double)(32 - __builtin_clz(mx));
    }
    case ORC_SR: return s_card + d_runs;
    case ORC_SW: return s_card + d_ewah;
    case ORC_RR: return ra + rb;
    case ORC_RW: return dmin(ra + wb, rb + wa);
    case ORC_WW: return wa + wb;
    default: return 0.0;
    }
}

double orc_predict(const orc_cost_model* m, int p, const orc_meta* a, const orc_meta* b,
                   double probe_scale, double margin, double rr_scale) {
    if (p == ORC_EMPTY) return 0.0;                      /* storm_cost.cpp:372-420 */
    if (p == ORC_BB && !(a->has_occ && b->has_occ))
        return m->ns_fixed + m->ns_per_unit[p] * (double)umax(a->n_words, b->n_words);
    double c = m->ns_fixed + m->ns_per_unit[p] * orc_work_units(p, a, b, margin) *
                                 (p == ORC_RR ? rr_scale : 1.0);
    if (p == ORC_BS && m->ns_probe_cold > m->ns_probe_hot) {
        const double touches = (a->cardinality <= b->cardinality) ? (double)a->n_nonzero_w
                                                                  : (double)b->n_nonzero_w;
        c += (m->ns_probe_cold - m->ns_probe_hot) * probe_scale * touches;
    }
    if (p == ORC_BB) {
        const double occ_words = dmax(1.0, (double)umax(a->n_words, b->n_words) / 512.0);
        c += m->ns_per_occ_word * occ_words;
    }
    return c;
}

int orc_select_pairing(const orc_cost_model* m, const orc_meta* a, const orc_meta* b,
                       double* predicted) {
    if (a->cardinality == 0 || b->cardinality == 0 ||      /* storm_cost.cpp:422-465 */
        a->last_set < b->first_set || b->last_set < a->first_set) {
        if (predicted) *predicted = 0.0;
        return ORC_EMPTY;
    }
    int best = ORC_BB;
    double bestc = orc_predict(m, best, a, b, 0.05, 16.0, 1.0);
    for (int i = 0; i < ORC_EMPTY; ++i) {
        if (m->ns_per_unit[i] <= 0.0) continue;
        const double c = orc_predict(m, i, a, b, 0.05, 16.0, 1.0);
        if (c < bestc) { bestc = c; best = i; }
    }
    if (predicted) *predicted = bestc;
    return best;
}

/* ------------------------------------------------------------------------ */
/* Driver helpers: kernels/storm_allpairs.cpp:281-323.                       */
/* ------------------------------------------------------------------------ */
void orc_tile_meta(const orc_meta* metas, const uint32_t* order, uint32_t lo, uint32_t hi,
                   orc_meta* m) {
    memset(m, 0, sizeof *m);
    m->first_set = UINT32_MAX;
    uint64_t card = 0, runs = 0, nnz = 0;
    const uint32_t cnt = hi > lo ? hi - lo : 1;
    m->has_rank = m->has_occ = (hi > lo);
    for (uint32_t k = lo; k < hi; ++k) {
        const orc_meta* r = &metas[order[k]];
        m->has_rank = m->has_rank && r->has_rank;
        m->has_occ = m->has_occ && r->has_occ;
        card += r->cardinality; runs += r->n_runs; nnz += r->n_nonzero_w;
        m->n_words = umax(m->n_words, r->n_words);
        if (r->cardinality) {
            if (r->first_set < m->first_set) m->first_set = r->first_set;
            if (r->last_set > m->last_set) m->last_set = r->last_set;
        }
    }
    m->cardinality = (uint32_t)(card / cnt);
    m->n_runs = (uint32_t)(runs / cnt);
    m->n_nonzero_w = (uint32_t)(nnz / cnt);
    if (m->first_set == UINT32_MAX) m->first_set = 0;
}

void orc_density_order(const orc_meta* metas, uint32_t n, uint32_t* order) {
    /* Stable insertion-free merge sort on (cardinality, index). */
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    uint32_t* tmp = (uint32_t*)malloc((size_t)(n ? n : 1) * 4);
    for (uint32_t w = 1; w < n; w *= 2) {
        for (uint32_t lo = 0; lo < n; lo += 2 * w) {
            uint32_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            uint32_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi)
                tmp[k++] = (metas[order[j]].cardinality < metas[order[i]].cardinality)
                               ? order[j++] : order[i++];
            while (i < mid) tmp[k++] = order[i++];
            while (j < hi) tmp[k++] = order[j++];
        }
        memcpy(order, tmp, (size_t)n * 4);
    }
    free(tmp);
}

/* ------------------------------------------------------------------------ */
/* Whole-matrix references.                                                  */
/* ------------------------------------------------------------------------ */
static inline uint64_t tri_index(uint64_t i, uint64_t j, uint64_t n) {
    return i * n - i * (i + 1) / 2 + (j - i - 1);
}

typedef struct {
    const uint32_t* offs; const uint32_t* pos; const uint64_t* words;
    uint32_t n, nw; uint32_t* out; int tid, nt; uint64_t sum;
} orc_job;

static void* orc_lists_worker(void* vp) {
    orc_job* j = (orc_job*)vp;
    uint64_t s = 0;
    for (uint32_t i = (uint32_t)j->tid; i < j->n; i += (uint32_t)j->nt)
        for (uint32_t k = i + 1; k < j->n; ++k) {
            const uint64_t c = orc_intersect_merge(j->pos + j->offs[i], j->offs[i + 1] - j->offs[i],
                                                   j->pos + j->offs[k], j->offs[k + 1] - j->offs[k]);
            if (j->out) j->out[tri_index(i, k, j->n)] = (uint32_t)c;
            s += c;
        }
    j->sum = s;
    return NULL;
}

static void* orc_words_worker(void* vp) {
    orc_job* j = (orc_job*)vp;
    uint64_t s = 0;
    for (uint32_t i = (uint32_t)j->tid; i < j->n; i += (uint32_t)j->nt)
        for (uint32_t k = i + 1; k < j->n; ++k) {
            const uint64_t* a = j->words + (size_t)i * j->nw;
            const uint64_t* b = j->words + (size_t)k * j->nw;
            uint64_t c = 0;
            for (uint32_t w = 0; w < j->nw; ++w) c += (uint64_t)__builtin_popcountll(a[w] & b[w]);
            if (j->out) j->out[tri_index(i, k, j->n)] = (uint32_t)c;
            s += c;
        }
    j->sum = s;
    return NULL;
}

static uint64_t orc_run_jobs(void* (*fn)(void*), orc_job proto, int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    orc_job jobs[256];
    pthread_t th[256];
    for (int t = 0; t < threads; ++t) {
        jobs[t] = proto; jobs[t].tid = t; jobs[t].nt = threads;
        if (threads > 1) pthread_create(&th[t], NULL, fn, &jobs[t]);
        else fn(&jobs[t]);
    }
    uint64_t s = 0;
    for (int t = 0; t < threads; ++t) {
        if (threads > 1) pthread_join(th[t], NULL);
        s += jobs[t].sum;
    }
    return s;
}

uint64_t orc_allpairs_lists(const uint32_t* offs, const uint32_t* pos, uint32_t n,
                            uint32_t* out_tri, int threads) {
    orc_job p; memset(&p, 0, sizeof p);
    p.offs = offs; p.pos = pos; p.n = n; p.out = out_tri;
    return orc_run_jobs(orc_lists_worker, p, threads);
}

uint64_t orc_allpairs_words(const uint64_t* words, uint32_t n, uint32_t nw,
                            uint32_t* out_tri, int threads) {
    orc_job p; memset(&p, 0, sizeof p);
    p.words = words; p.n = n; p.nw = nw; p.out = out_tri;
    return orc_run_jobs(orc_words_worker, p, threads);
}

/* Rows [r0, r1) of the packed triangle: for each such i every j in (i, n), in
 * triangle order (one contiguous range of it). Threads split each row's j. */
typedef struct {
    const uint64_t* words; uint32_t n, nw, r0, r1; uint32_t* out; int tid, nt; uint64_t sum;
} orc_band_job;

static void* orc_band_worker(void* vp) {
    orc_band_job* j = (orc_band_job*)vp;
    uint64_t s = 0;
    const uint64_t base = tri_index(j->r0, j->r0 + 1, j->n);
    for (uint32_t i = j->r0; i < j->r1; ++i) {
        const uint64_t* a = j->words + (size_t)i * j->nw;
        for (uint32_t k = i + 1 + (uint32_t)j->tid; k < j->n; k += (uint32_t)j->nt) {
            const uint64_t* b = j->words + (size_t)k * j->nw;
            uint64_t c = 0;
            for (uint32_t w = 0; w < j->nw; ++w) c += (uint64_t)__builtin_popcountll(a[w] & b[w]);
            if (j->out) j->out[tri_index(i, k, j->n) - base] = (uint32_t)c;
            s += c;
        }
    }
    j->sum = s;
    return NULL;
}

uint64_t orc_band_words(const uint64_t* words, uint32_t n, uint32_t nw, uint32_t r0, uint32_t r1,
                        uint32_t* out, int threads) {
    if (r1 > n) r1 = n;
    if (r0 >= r1 || r0 + 1 >= n) return 0;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    orc_band_job jobs[256];
    pthread_t th[256];
    for (int t = 0; t < threads; ++t) {
        orc_band_job b = {words, n, nw, r0, r1, out, t, threads, 0};
        jobs[t] = b;
        if (threads > 1) pthread_create(&th[t], NULL, orc_band_worker, &jobs[t]);
        else orc_band_worker(&jobs[t]);
    }
    uint64_t s = 0;
    for (int t = 0; t < threads; ++t) {
        if (threads > 1) pthread_join(th[t], NULL);
        s += jobs[t].sum;
    }
    return s;
}

/* Column counts and row sums of a dense matrix on `threads` threads: column
 * words split between threads for the counts, rows for the sums. */
typedef struct {
    const uint64_t* words; uint32_t n, nw; uint32_t* col; uint64_t* row_sum; int tid, nt, phase;
} orc_cols_job;

static void* orc_cols_worker(void* vp) {
    orc_cols_job* j = (orc_cols_job*)vp;
    if (j->phase == 0) {
        const uint32_t w0 = (uint32_t)((uint64_t)j->nw * j->tid / j->nt);
        const uint32_t w1 = (uint32_t)((uint64_t)j->nw * (j->tid + 1) / j->nt);
        memset(j->col + (size_t)w0 * 64, 0, (size_t)(w1 - w0) * 64 * 4);
        for (uint32_t i = 0; i < j->n; ++i)
            for (uint32_t w = w0; w < w1; ++w) {
                uint64_t x = j->words[(size_t)i * j->nw + w];
                while (x) { j->col[(size_t)w * 64 + __builtin_ctzll(x)]++; x &= x - 1; }
            }
    } else {
        for (uint32_t i = (uint32_t)j->tid; i < j->n; i += (uint32_t)j->nt) {
            uint64_t s = 0;
            for (uint32_t w = 0; w < j->nw; ++w) {
                uint64_t x = j->words[(size_t)i * j->nw + w];
                while (x) { s += j->col[(size_t)w * 64 + __builtin_ctzll(x)] - 1; x &= x - 1; }
            }
            j->row_sum[i] = s;
        }
    }
    return NULL;
}

void orc_column_closed_forms_words(const uint64_t* words, uint32_t n, uint32_t nw, uint32_t* col,
                                   uint64_t* row_sum, int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    orc_cols_job jobs[256];
    pthread_t th[256];
    for (int phase = 0; phase < 2; ++phase) {
        if (phase == 1 && !row_sum) break;
        for (int t = 0; t < threads; ++t) {
            orc_cols_job b = {words, n, nw, col, row_sum, t, threads, phase};
            jobs[t] = b;
            if (threads > 1) pthread_create(&th[t], NULL, orc_cols_worker, &jobs[t]);
            else orc_cols_worker(&jobs[t]);
        }
        for (int t = 0; t < threads && threads > 1; ++t) pthread_join(th[t], NULL);
    }
}

void orc_column_counts_words(const uint64_t* words, uint32_t n, uint32_t nw, uint32_t* col) {
    memset(col, 0, (size_t)nw * 64 * 4);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t w = 0; w < nw; ++w) {
            uint64_t x = words[(size_t)i * nw + w];
            while (x) { col[(size_t)w * 64 + __builtin_ctzll(x)]++; x &= x - 1; }
        }
}

uint64_t orc_total_from_columns(const uint32_t* col, uint64_t n_cols) {
    uint64_t t = 0;
    for (uint64_t b = 0; b < n_cols; ++b) t += (uint64_t)col[b] * (col[b] ? col[b] - 1 : 0) / 2;
    return t;
}

void orc_row_sums_words(const uint64_t* words, uint32_t n, uint32_t nw, const uint32_t* col,
                        uint64_t* row_sum) {
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t s = 0;
        for (uint32_t w = 0; w < nw; ++w) {
            uint64_t x = words[(size_t)i * nw + w];
            while (x) { s += col[(size_t)w * 64 + __builtin_ctzll(x)] - 1; x &= x - 1; }
        }
        row_sum[i] = s;
    }
}

void orc_row_sums_from_tri(const uint32_t* tri, uint32_t n, uint64_t* row_sum) {
    memset(row_sum, 0, (size_t)n * 8);
    uint64_t idx = 0;
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = i + 1; j < n; ++j, ++idx) {
            row_sum[i] += tri[idx];
            row_sum[j] += tri[idx];
        }
}

/* ------------------------------------------------------------------------ */
/* Learned-selector math in float64. No reference counterpart: pinned by     */
/* finite differences and by the published Adam update (Kingma & Ba 2015).  */
/* ------------------------------------------------------------------------ */
double orc_mlp_forward_backward(int f, int h1, int h2, int o, int batch,
                                const double* x, const double* t, const uint8_t* mask,
                                const double* W1, const double* b1,
                                const double* W2, const double* b2,
                                const double* W3, const double* b3,
                                double* y,
                                double* gW1, double* gb1, double* gW2, double* gb2,
                                double* gW3, double* gb3) {
    double* a1 = (double*)malloc(sizeof(double) * h1);
    double* a2 = (double*)malloc(sizeof(double) * h2);
    double* d1 = (double*)malloc(sizeof(double) * h1);
    double* d2 = (double*)malloc(sizeof(double) * h2);
    double* yo = (double*)malloc(sizeof(double) * o);
    double* dy = (double*)malloc(sizeof(double) * o);
    if (gW1) {
        memset(gW1, 0, sizeof(double) * h1 * f); memset(gb1, 0, sizeof(double) * h1);
        memset(gW2, 0, sizeof(double) * h2 * h1); memset(gb2, 0, sizeof(double) * h2);
        memset(gW3, 0, sizeof(double) * o * h2); memset(gb3, 0, sizeof(double) * o);
    }
    uint64_t cnt = 0;
    for (int n = 0; n < batch; ++n)
        for (int k = 0; k < o; ++k) cnt += mask[(size_t)n * o + k] ? 1 : 0;
    const double inv = cnt ? 1.0 / (double)cnt : 0.0;
    double loss = 0.0;
    for (int n = 0; n < batch; ++n) {
        const double* xn = x + (size_t)n * f;
        for (int i = 0; i < h1; ++i) {
            double s = b1[i];
            for (int k = 0; k < f; ++k) s += W1[(size_t)i * f + k] * xn[k];
            a1[i] = s > 0 ? s : 0;
        }
        for (int i = 0; i < h2; ++i) {
            double s = b2[i];
            for (int k = 0; k < h1; ++k) s += W2[(size_t)i * h1 + k] * a1[k];
            a2[i] = s > 0 ? s : 0;
        }
        for (int i = 0; i < o; ++i) {
            double s = b3[i];
            for (int k = 0; k < h2; ++k) s += W3[(size_t)i * h2 + k] * a2[k];
            yo[i] = s;
            if (y) y[(size_t)n * o + i] = s;
            const int mk = mask[(size_t)n * o + i] != 0;
            const double e = mk ? (s - t[(size_t)n * o + i]) : 0.0;
            loss += 0.5 * e * e * inv;
            dy[i] = e * inv;
        }
        if (!gW1) continue;
        for (int i = 0; i < o; ++i) {
            gb3[i] += dy[i];
            for (int k = 0; k < h2; ++k) gW3[(size_t)i * h2 + k] += dy[i] * a2[k];
        }
        for (int k = 0; k < h2; ++k) {
            double s = 0;
            for (int i = 0; i < o; ++i) s += W3[(size_t)i * h2 + k] * dy[i];
            d2[k] = a2[k] > 0 ? s : 0;
        }
        for (int i = 0; i < h2; ++i) {
            gb2[i] += d2[i];
            for (int k = 0; k < h1; ++k) gW2[(size_t)i * h1 + k] += d2[i] * a1[k];
        }
        for (int k = 0; k < h1; ++k) {
            double s = 0;
            for (int i = 0; i < h2; ++i) s += W2[(size_t)i * h1 + k] * d2[i];
            d1[k] = a1[k] > 0 ? s : 0;
        }
        for (int i = 0; i < h1; ++i) {
            gb1[i] += d1[i];
            for (int k = 0; k < f; ++k) gW1[(size_t)i * f + k] += d1[i] * xn[k];
        }
    }
    free(a1); free(a2); free(d1); free(d2); free(yo); free(dy);
    return loss;
}

void orc_adam_step(int n, double* p, const double* g, double* m, double* v, int t,
                   double lr, double beta1, double beta2, double eps) {
    const double c1 = 1.0 - pow(beta1, t), c2 = 1.0 - pow(beta2, t);
    for (int i = 0; i < n; ++i) {
        m[i] = beta1 * m[i] + (1.0 - beta1) * g[i];
        v[i] = beta2 * v[i] + (1.0 - beta2) * g[i] * g[i];
        const double mh = m[i] / c1, vh = v[i] / c2;
        p[i] -= lr * mh / (sqrt(vh) + eps);
    }
}

/* ------------------------------------------------------------------------ */
/* Host ISA check (which reference build may run here).                      */
/* ------------------------------------------------------------------------ */
int orc_cpu_has_avx512_vpopcntdq(void) {
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
           __builtin_cpu_supports("avx512vl") && __builtin_cpu_supports("avx512vpopcntdq");
}
