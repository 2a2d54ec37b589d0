// Internal definitions shared by the engine's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "storm_hip.h"

namespace stormhip {

// Row padding of dense corpora: rows are padded to a multiple of this many
// 64-bit words (zero filled) so every B x B K-stage is a full, aligned load.
constexpr uint32_t kWordPad = 64;
// Row-count padding: dense corpora hold a multiple of this many rows (zero
// rows past n_rows), so a B x B tile never reads outside the allocation.
constexpr uint32_t kRowPad = 256;

void set_error(const char* fmt, ...);

#define SH_HIP(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            ::stormhip::set_error("%s failed: %s (%s:%d)", #call,             \
                                  hipGetErrorString(e_), __FILE__, __LINE__); \
            return e_ == hipErrorOutOfMemory ? STORM_HIP_ERR_NOMEM            \
                                             : STORM_HIP_ERR_DEVICE;          \
        }                                                                     \
    } while (0)

#define SH_TRY(expr)                        \
    do {                                    \
        int s_ = (expr);                    \
        if (s_ != STORM_HIP_OK) return s_;  \
    } while (0)

struct DeviceCostModel {
    double ns_per_unit[10];
    double ns_fixed, ns_per_occ_word, ns_probe_hot, ns_probe_cold;
    double probe_scale, br_margin, rr_scale;
    double bb_zonemap;
};

// Launch shape of the matrix-core B x B kernel (chosen by the autotuner).
struct MfmaConfig {
    uint32_t stage_words = 16;    // packed words per LDS stage: 4, 8 or 16
    uint32_t band = 4;            // tile order: 0 row-major, else bands of this many tile rows
    uint32_t blocks_per_cu = 2;   // K-split target when there are fewer tiles than this per CU
    uint32_t waves = 8;           // 8 (two per SIMD, 128 x 64 each) or 4 (one per SIMD, 128 x 128)
};

}  // namespace stormhip

struct storm_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // Side streams (default priority) so cell classes run concurrently:
    // forked from and joined back into `stream` with the fork/join events.
    hipStream_t aux[2] = {};
    hipEvent_t fork_ev = nullptr, join_ev[2] = {};
    hipEvent_t ev[8] = {};
    stormhip::DeviceCostModel model{};
    stormhip::MfmaConfig mfma{};     // B x B launch shape (defaults or tuned)
    int cu_count = 0;
    // MLP selector state lives in mlp.hip (opaque here).
    void* selector = nullptr;
    // Bumped whenever the cost model or the selector changes: captured
    // graphs bake both in and are rebuilt when it moves.
    uint32_t model_version = 1;
};

struct storm_hip_corpus {
    storm_hip_ctx* ctx = nullptr;
    uint32_t n_rows = 0;
    uint32_t n_words = 0;      // logical words per row (ceil(universe/64))
    uint32_t universe = 0;     // bits
    bool dense = false;        // contiguous corpus: every row a bitmap
    // Representations held for every row (bitmaps are per row, below). A
    // corpus without lists holds its rows as runs or as bitmaps only; the
    // cells that read a list are then not eligible.
    bool has_lists = true;
    bool has_runs = false, has_ewah = false, has_comp = false;
    // Bit-sliced bitmaps (STORM_HIP_BUILD_SLICES): for each group of 64
    // density-ordered rows that all hold bitmaps, per position p a mask whose
    // bit k says row 64g + k holds p; the groups of 512 consecutive rows share
    // a slab of ld * 64 positions x 8 masks. d_slice_of[g] = 8 * slab + g % 8
    // or kNoBitmap.
    bool want_slices = false;
    uint64_t* d_slice = nullptr;
    uint32_t* d_slice_of = nullptr;
    uint32_t n_slices = 0;
    // Coarse zone maps (STORM_HIP_BUILD_ZONES): row r's map is zone_ld words
    // at d_zone + r * zone_ld, bit b set when the row holds a position p with
    // p >> zone_shift == b.
    bool want_zones = false;
    uint64_t* d_zone = nullptr;
    uint32_t zone_ld = 0, zone_shift = 0;
    // Dense layout: n_rows_pad x ld words.
    uint64_t* d_words = nullptr;
    uint32_t ld = 0;
    uint32_t n_rows_pad = 0;
    // General layout (arenas). Offsets are u64 element offsets, n_rows+1.
    uint64_t *d_list_off = nullptr, *d_run_off = nullptr, *d_ewah_off = nullptr;
    uint64_t* d_comp_off = nullptr;
    uint32_t *d_list = nullptr, *d_run_start = nullptr, *d_run_end = nullptr;
    uint64_t *d_ewah = nullptr, *d_bitmap = nullptr, *d_occ = nullptr;
    uint32_t* d_comp = nullptr;
    uint32_t* d_meta = nullptr;      // 8 per row
    // EWAH marker tables (decoded start word and buffer offset per marker).
    uint64_t* d_mk_off = nullptr;
    uint32_t *d_mk_word = nullptr, *d_mk_pos = nullptr;
    // Bitmap store: row r's bitmap is d_bitmap + d_bm_row[r] * ld, or absent
    // (kNoBitmap). One extra all-zero row sits at index bm_capacity.
    uint32_t* d_bm_row = nullptr;
    uint32_t n_bitmap_rows = 0, bm_capacity = 0;
    std::vector<uint32_t> h_bm_row;
    // The store already holds every bitmap row (generated there): finishing
    // the corpus keeps it instead of filling it from lists or runs.
    bool store_ready = false;
    // rank9 (kernels/storm_repr.cpp:25-45) and zone maps (:47-59) of the
    // bitmap rows whose build_row gates pass (meta fields 6 and 7): row r's
    // index is d_rank + d_rank_row[r] * rank_ld (kNoBitmap: none), likewise
    // d_occ / d_occ_row / occ_ld.
    uint64_t* d_rank = nullptr;
    uint32_t *d_rank_row = nullptr, *d_occ_row = nullptr;
    uint32_t rank_ld = 0, occ_ld = 0, n_rank_rows = 0, n_occ_rows = 0;
    uint64_t index_bytes = 0;
    // Density order (stable by cardinality) and the identity order, both
    // padded with the zero-row sentinel mapping; d_inv: row -> its position
    // in the density order.
    uint32_t *d_perm = nullptr, *d_ident = nullptr, *d_inv = nullptr;
    // Per density position: the row's cardinality, first and last position
    // (4 u32 each, padding positions empty).
    uint32_t* d_pspan = nullptr;
    std::vector<uint32_t> h_perm;
    uint64_t device_bytes = 0;
    std::vector<uint32_t> h_meta;    // host copy of meta (8 per row)
    bool bitmaps_complete = false;
    void* work = nullptr;            // per-corpus selection state of the general path
    void* plan = nullptr;            // cached launch plan of its single-corpus calls
    // Result of the last UPPER_TRI / UPPER_TRI_DENSITY call (tri_density:
    // packed over density positions).
    uint32_t* d_tri = nullptr;
    uint64_t tri_count = 0;
    bool tri_valid = false;
    bool tri_density = false;
    // Decisions of the last selecting call.
    std::vector<uint8_t> last_cells;
    // Thresholded calls: device pair buffer (capacity in pairs) and counter,
    // and the last call's pairs sorted by (i, j), 3 u32 each.
    void* d_hits = nullptr;
    uint64_t hit_cap = 0;
    // Band and cross calls: device count buffer, sized per call shape.
    uint32_t* d_emit = nullptr;
    uint64_t emit_bytes = 0;
    unsigned long long* d_n_hits = nullptr;
    std::vector<uint32_t> h_hits;
    bool hits_valid = false;
};

namespace stormhip {
// Looks up the allocation-free call's scratch array on this (current)
// device, once per device (noalloc.hip).
int noalloc_register_device(int device);
// rank9 and zone maps of the bitmap rows whose gates pass (gen_dev.hip).
int build_indexes(storm_hip_corpus* c, hipStream_t s);
// Fill the bitmap rows of a corpus held as runs (gen_dev.hip).
int fill_bitmaps_from_runs(storm_hip_corpus* c, hipStream_t s);
// Grid launches over all upper-triangle tiles of one bitmap matrix.
int launch_bb_valu(hipStream_t s, int bm, const uint64_t* X, uint32_t ld, uint32_t n_words,
                   const uint32_t* rowmap, const uint32_t* perm, uint32_t n, uint32_t* tri,
                   unsigned long long* sum);
int launch_bb_mfma_fmt(bool fp4, const MfmaConfig& cfg, hipStream_t s, const uint64_t* X,
                       uint32_t ld, uint32_t n_words, const uint32_t* rowmap, const uint32_t* perm,
                       uint32_t n, uint32_t* tri, unsigned long long* sum, uint32_t k_split = 1,
                       uint32_t* partial = nullptr);
// Queue launches over tiles of several corpora (declared with the device
// corpus view in general.h).
// The B x B form for a dense launch: VALU tile of 128, or MFMA (int8 / fp4).
// AUTO takes fp4 where its exactness bound holds, else int8. Small corpora
// split K across blocks so the launch fills `cu_count` CUs.
// perm (optional): the triangle is written for rows perm[i], perm[j] of the
// walk's positions i, j (rowmap gives their bitmap rows). may_allocate: a
// K-split launch may take a stream-ordered scratch buffer for the slices'
// partial counts (otherwise the slices add into the zeroed triangle).
int launch_bb_dense(uint32_t bb_kernel, const MfmaConfig& cfg, hipStream_t s, const uint64_t* X,
                    uint32_t ld, uint32_t n_words, const uint32_t* rowmap, uint32_t n,
                    uint32_t* tri, unsigned long long* sum, uint32_t cu_count,
                    const uint32_t* perm = nullptr, bool may_allocate = true);
// Rectangular X . Y^T on the matrix cores (rows of both padded to 256, same
// stride); out: nx * ny u32 row-major, or null for the sum only.
int launch_bb_rect(bool fp4, const MfmaConfig& cfg, hipStream_t s, const uint64_t* X, uint32_t nx,
                   const uint64_t* Y, uint32_t ny, uint32_t ld, uint32_t n_words, uint32_t* out,
                   unsigned long long* sum, uint32_t cu_count);
}  // namespace stormhip
