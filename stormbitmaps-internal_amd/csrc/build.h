// Corpus construction shared by the host builder (build.hip) and the device
// builder (build_dev.hip).
#pragma once

#include <cstdint>
#include <vector>

#include "internal.h"

namespace stormhip {

// A general corpus with its scalar fields set and no device arrays yet.
storm_hip_corpus* corpus_new(storm_hip_ctx* ctx, uint32_t n, uint32_t universe, uint32_t flags);

// Rows that get a bitmap under a placement policy (flags bits 0-1), from the
// per-row meta (cardinality is field 0).
std::vector<uint8_t> bitmap_rows(const std::vector<uint32_t>& meta, uint32_t n, uint32_t universe,
                                 uint32_t policy);

// Allocates and uploads a host vector (at least one element is allocated).
template <class T>
int upload_vec(T*& d, const std::vector<T>& h, hipStream_t s, uint64_t& bytes) {
    const size_t b = (h.empty() ? 1 : h.size()) * sizeof(T);
    SH_HIP(hipMalloc((void**)&d, b));
    bytes += b;
    if (!h.empty())
        SH_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return STORM_HIP_OK;
}

// Density order, identity order, bitmap row map and bitmap store of a corpus
// whose arenas and host meta are in place; synchronises the stream.
int corpus_finish(storm_hip_ctx* ctx, storm_hip_corpus* c, const std::vector<uint8_t>& want_bitmap);
// The same for a corpus whose bitmap store is already complete (one slot per
// row, slot = row): orders, the store rearranged into density order (bm_row
// updated), indexes and slices.
int corpus_finish_store(storm_hip_ctx* ctx, storm_hip_corpus* c);

// Builds runs, EWAH (+ marker tables), complements and meta on the device
// from the lists already at c->d_list_off / c->d_list (flags: the
// STORM_HIP_BUILD_* bits). Fills c->h_meta. STORM_HIP_ERR_INVALID when a row
// is not sorted, distinct and inside the universe.
int device_build(storm_hip_ctx* ctx, storm_hip_corpus* c, uint32_t flags);

}  // namespace stormhip
