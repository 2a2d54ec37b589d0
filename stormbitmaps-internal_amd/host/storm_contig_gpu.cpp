// STORM_contiguous_t on the GPU: the host keeps the rows as bitmap words (the
// layout of storm.cpp:1113-1157) and the pairwise calls upload them once per
// change and run the all-bitmap B x B cell on the device.
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "storm_gpu.h"
#include "storm_hip.h"

struct STORM_contiguous_s {
    uint64_t vector_length = 0;
    uint32_t n_words = 0;
    std::vector<uint64_t> data;   // n_words per vector
    size_t n = 0;
    storm_hip_ctx* ctx = nullptr;
    storm_hip_corpus* corpus = nullptr;
    bool dirty = true;
    bool counts_valid = false;
};

namespace {

void drop_corpus(STORM_contiguous_t* h) {
    if (h->corpus) storm_hip_corpus_free(h->corpus);
    h->corpus = nullptr;
    h->dirty = true;
    h->counts_valid = false;
}

int ensure_device(STORM_contiguous_t* h) {
    if (!h->ctx) {
        const int s = storm_hip_ctx_create(0, 0, &h->ctx);
        if (s != STORM_HIP_OK) return s;
    }
    if (h->dirty || !h->corpus) {
        if (h->corpus) storm_hip_corpus_free(h->corpus);
        h->corpus = nullptr;
        const int s = storm_hip_corpus_upload_contig(h->ctx, h->data.data(), (uint32_t)h->n,
                                                     h->n_words, &h->corpus);
        if (s != STORM_HIP_OK) return s;
        h->dirty = false;
    }
    return STORM_HIP_OK;
}

uint64_t pairw(STORM_contiguous_t* h, uint32_t out_mode) {
    if (h == nullptr) return 0;
    if (h->n < 2) return 0;
    int s = ensure_device(h);
    storm_hip_stats st;
    if (s == STORM_HIP_OK) {
        storm_hip_options o;
        std::memset(&o, 0, sizeof o);
        o.policy = STORM_HIP_POLICY_ALL_BITMAP;
        o.out_mode = out_mode;
        o.bb_kernel = STORM_HIP_BB_AUTO;
        s = storm_hip_allpairs(h->ctx, h->corpus, &o, &st);
    }
    if (s != STORM_HIP_OK) {
        std::fprintf(stderr, "STORM_contig (GPU): %s: %s\n", storm_hip_strerror(s),
                     storm_hip_last_error());
        return UINT64_MAX;
    }
    h->counts_valid = (out_mode == STORM_HIP_OUT_UPPER_TRI);
    return st.sum;
}

}  // namespace

extern "C" {

STORM_contiguous_t* STORM_contig_new(size_t vector_length) {
    STORM_contiguous_t* h = new (std::nothrow) STORM_contiguous_t();
    if (!h) return nullptr;
    h->vector_length = vector_length;
    h->n_words = (uint32_t)((vector_length + 63) / 64);
    return h;
}

void STORM_contig_free(STORM_contiguous_t* h) {
    if (!h) return;
    drop_corpus(h);
    if (h->ctx) storm_hip_ctx_free(h->ctx);
    delete h;
}

int STORM_contig_add(STORM_contiguous_t* h, const uint32_t* values, const uint32_t n_values) {
    if (h == nullptr) return -1;
    if (values == nullptr) return -2;
    if (n_values == 0) return 0;
    for (uint32_t i = 0; i < n_values; ++i) {
        if (values[i] >= h->vector_length) return -3;
    }
    try {
        const size_t base = h->data.size();
        h->data.resize(base + h->n_words, 0);
        uint64_t* dst = h->data.data() + base;
        for (uint32_t i = 0; i < n_values; ++i) dst[values[i] >> 6] |= 1ull << (values[i] & 63);
    } catch (...) {
        return -4;
    }
    ++h->n;
    h->dirty = true;
    h->counts_valid = false;
    return (int)n_values;
}

int STORM_contig_clear(STORM_contiguous_t* h) {
    if (h == nullptr) return -1;
    h->data.clear();
    h->n = 0;
    drop_corpus(h);
    return 1;
}

uint64_t STORM_contig_pairw_intersect_cardinality(STORM_contiguous_t* h) {
    return pairw(h, STORM_HIP_OUT_SUM);
}
uint64_t STORM_contig_pairw_intersect_cardinality_blocked(STORM_contiguous_t* h, uint32_t) {
    return pairw(h, STORM_HIP_OUT_SUM);
}
uint64_t STORM_contig_pairw_intersect_cardinality_list(STORM_contiguous_t* h) {
    return pairw(h, STORM_HIP_OUT_SUM);
}
uint64_t STORM_contig_pairw_intersect_cardinality_blocked_list(STORM_contiguous_t* h, uint32_t) {
    return pairw(h, STORM_HIP_OUT_SUM);
}

int STORM_contig_gpu_pairw_counts(STORM_contiguous_t* h, uint32_t* out) {
    if (!h) return STORM_HIP_ERR_INVALID;
    if (h->n < 2) return STORM_HIP_OK;
    if (!out) return STORM_HIP_ERR_NULL_VALUES;
    if (pairw(h, STORM_HIP_OUT_UPPER_TRI) == UINT64_MAX) return STORM_HIP_ERR_DEVICE;
    const uint64_t n = h->n;
    return storm_hip_result_copy_tri(h->ctx, h->corpus, 0, n * (n - 1) / 2, out);
}

size_t STORM_contig_gpu_size(const STORM_contiguous_t* h) { return h ? h->n : 0; }

}  // extern "C"
