// Host-side representation building and device upload of general corpora.
//
// From sorted, distinct per-row positions this builds what build_row
// (kernels/storm_repr.cpp:102-207) builds -- runs, EWAH-64, complement and the
// RowMeta fields with the same index gates -- in O(card) per row (the
// reference scans the whole universe per row), multi-threaded over rows, then
// uploads the arenas. Bitmaps are scattered on the device from the lists.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <new>
#include <numeric>
#include <thread>
#include <vector>

#include "device_common.h"
#include "build.h"
#include "general.h"
#include "internal.h"

namespace stormhip {

namespace {

struct HostArena {
    uint32_t n = 0, universe = 0, nw = 0;
    bool has_runs = false, has_ewah = false, has_comp = false;
    std::vector<uint64_t> list_off, run_off, ewah_off, mk_off, comp_off;
    std::vector<uint32_t> list, run_start, run_end, mk_word, mk_pos, comp, meta;
    std::vector<uint64_t> ewah;
    std::vector<uint8_t> want_bitmap;
};

// Nonzero words of a sorted position list, in order.
struct WordWalker {
    const uint32_t* p; uint32_t n, i = 0;
    WordWalker(const uint32_t* p_, uint32_t n_) : p(p_), n(n_) {}
    bool next(uint32_t& w, uint64_t& v) {
        if (i >= n) return false;
        w = p[i] >> 6;
        v = 0;
        while (i < n && (p[i] >> 6) == w) { v |= 1ull << (p[i] & 63); ++i; }
        return true;
    }
};

// EWAH-64 of the row (build_ewah, kernels/storm_repr.cpp:78-100) from its
// nonzero words; when `out` is null only sizes are returned.
void ewah_from_list(const uint32_t* p, uint32_t n, uint32_t nw, uint64_t* out, uint32_t* mk_word,
                    uint32_t* mk_pos, uint32_t& n_out, uint32_t& n_mk) {
    // Nonzero words of the row.
    std::vector<uint32_t> wi;
    std::vector<uint64_t> wv;
    {
        WordWalker ww(p, n);
        uint32_t w; uint64_t v;
        while (ww.next(w, v)) { wi.push_back(w); wv.push_back(v); }
    }
    size_t k = 0;               // next nonzero word
    uint32_t i = 0, o = 0, m = 0;
    auto word_at = [&](uint32_t x) -> uint64_t {
        return (k < wi.size() && wi[k] == x) ? wv[k] : 0ull;
    };
    while (i < nw) {
        uint64_t cur = word_at(i);
        bool fv = (cur == ~0ull);
        uint64_t fl = 0;
        const uint32_t start = i;
        if (cur == 0) {
            const uint32_t stop = (k < wi.size()) ? wi[k] : nw;
            const uint64_t len = std::min<uint64_t>(stop - i, 0x7FFFFFFFull);
            fl = len; i += (uint32_t)len;
        } else if (fv) {
            while (i < nw && k < wi.size() && wi[k] == i && wv[k] == ~0ull && fl < 0x7FFFFFFFull) {
                ++i; ++k; ++fl;
            }
        } else {
            fv = false;
        }
        const uint32_t lit0 = i;
        const size_t k0 = k;
        while (i < nw && k < wi.size() && wi[k] == i && wv[k] != ~0ull) { ++i; ++k; }
        const uint64_t ll = i - lit0;
        if (out) {
            mk_word[m] = start;
            mk_pos[m] = o;
            out[o] = (uint64_t)fv | (fl << 1) | (ll << 32);
            for (uint64_t q = 0; q < ll; ++q) out[o + 1 + q] = wv[k0 + q];
        }
        o += 1 + (uint32_t)ll;
        ++m;
    }
    if (o == 0) {   // empty universe: one empty marker
        if (out) { out[0] = 0; mk_word[0] = 0; mk_pos[0] = 0; }
        o = 1; m = 1;
    }
    n_out = o;
    n_mk = m;
}

void parallel_rows(uint32_t n, const std::function<void(uint32_t)>& fn) {
    unsigned th = std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
    if (n < 64) th = 1;
    std::atomic<uint32_t> next{0};
    auto work = [&] {
        for (;;) {
            const uint32_t lo = next.fetch_add(64);
            if (lo >= n) break;
            for (uint32_t r = lo; r < std::min(n, lo + 64); ++r) fn(r);
        }
    };
    std::vector<std::thread> ts;
    for (unsigned t = 1; t < th; ++t) ts.emplace_back(work);
    work();
    for (auto& t : ts) t.join();
}

int build_host_arena(const uint64_t* offs, const uint32_t* pos, uint32_t n, uint32_t universe,
                     uint32_t flags, HostArena& A) {
    const uint32_t bitmap_policy = flags & 3u;
    A.n = n; A.universe = universe; A.nw = (universe + 63) / 64;
    A.has_runs = flags & STORM_HIP_BUILD_RUNS;
    A.has_ewah = flags & STORM_HIP_BUILD_EWAH;
    A.has_comp = flags & STORM_HIP_BUILD_COMP;
    // Validate: sorted, distinct, in range.
    std::atomic<int> bad{0};
    parallel_rows(n, [&](uint32_t r) {
        for (uint64_t i = offs[r]; i < offs[r + 1]; ++i)
            if (pos[i] >= universe || (i > offs[r] && pos[i] <= pos[i - 1])) { bad = 1; return; }
    });
    if (bad) { set_error("row positions must be sorted, distinct and < universe"); return STORM_HIP_ERR_INVALID; }

    std::vector<uint32_t> nrun(n), newah(n), nmk(n), ncomp(n);
    A.meta.assign((size_t)n * 8, 0);
    A.want_bitmap.assign(n, 0);
    parallel_rows(n, [&](uint32_t r) {
        const uint32_t* p = pos + offs[r];
        const uint32_t c = (uint32_t)(offs[r + 1] - offs[r]);
        uint32_t runs = 0, nnz = 0, lastw = 0xFFFFFFFFu;
        for (uint32_t i = 0; i < c; ++i) {
            if (i == 0 || p[i] != p[i - 1] + 1) ++runs;
            if ((p[i] >> 6) != lastw) { ++nnz; lastw = p[i] >> 6; }
        }
        nrun[r] = runs;
        uint32_t* m = &A.meta[(size_t)r * 8];
        m[0] = c; m[1] = runs; m[2] = A.nw; m[3] = nnz;
        m[4] = c ? p[0] : 0xFFFFFFFFu; m[5] = c ? p[c - 1] : 0;
        const uint64_t mean_rl = runs ? (uint64_t)c / runs : 0;
        m[6] = mean_rl >= 512;                                  // rank gate
        m[7] = (uint64_t)c * 512 > (uint64_t)universe;          // occ gate
        A.want_bitmap[r] = bitmap_policy == 1 ||
                           (bitmap_policy == 0 && (uint64_t)c * 64 >= (uint64_t)universe);
        if (A.has_ewah) {
            uint32_t o, mk;
            ewah_from_list(p, c, A.nw, nullptr, nullptr, nullptr, o, mk);
            newah[r] = o; nmk[r] = mk;
        }
        ncomp[r] = (A.has_comp && (uint64_t)c * 2 > universe) ? universe - c : 0;
    });
    auto prefix = [n](const std::vector<uint32_t>& cnt, std::vector<uint64_t>& off) {
        off.assign((size_t)n + 1, 0);
        for (uint32_t r = 0; r < n; ++r) off[r + 1] = off[r] + cnt[r];
    };
    A.list_off.assign(offs, offs + n + 1);
    A.list.assign(pos + offs[0], pos + offs[n]);
    if (offs[0] != 0) for (auto& x : A.list_off) x -= offs[0];
    try {
        if (A.has_runs) {
            prefix(nrun, A.run_off);
            A.run_start.resize(A.run_off[n]);
            A.run_end.resize(A.run_off[n]);
        } else {
            A.run_off.assign((size_t)n + 1, 0);
        }
        if (A.has_ewah) {
            prefix(newah, A.ewah_off);
            prefix(nmk, A.mk_off);
            A.ewah.resize(A.ewah_off[n]);
            A.mk_word.resize(A.mk_off[n]);
            A.mk_pos.resize(A.mk_off[n]);
        } else {
            A.ewah_off.assign((size_t)n + 1, 0);
            A.mk_off.assign((size_t)n + 1, 0);
        }
        prefix(ncomp, A.comp_off);
        A.comp.resize(A.comp_off[n]);
    } catch (...) {
        set_error("host allocation failed building arenas");
        return STORM_HIP_ERR_NOMEM;
    }
    parallel_rows(n, [&](uint32_t r) {
        const uint32_t* p = pos + offs[r];
        const uint32_t c = (uint32_t)(offs[r + 1] - offs[r]);
        if (A.has_runs) {
            uint64_t k = A.run_off[r];
            for (uint32_t i = 0; i < c;) {
                uint32_t j = i;
                while (j + 1 < c && p[j + 1] == p[j] + 1) ++j;
                A.run_start[k] = p[i]; A.run_end[k] = p[j] + 1; ++k;
                i = j + 1;
            }
        }
        if (A.has_ewah) {
            uint32_t o, mk;
            ewah_from_list(p, c, A.nw, &A.ewah[A.ewah_off[r]], &A.mk_word[A.mk_off[r]],
                           &A.mk_pos[A.mk_off[r]], o, mk);
        }
        if (A.comp_off[r + 1] > A.comp_off[r]) {
            uint64_t k = A.comp_off[r];
            uint32_t prev = 0;
            for (uint32_t i = 0; i < c; ++i) {
                for (uint32_t q = prev; q < p[i]; ++q) A.comp[k++] = q;
                prev = p[i] + 1;
            }
            for (uint32_t q = prev; q < universe; ++q) A.comp[k++] = q;
        }
    });
    return STORM_HIP_OK;
}

// Scatter the list of every row that has a bitmap into its bitmap row.
__global__ void scatter_bits_kernel(const uint64_t* __restrict__ list_off,
                                    const uint32_t* __restrict__ list,
                                    const uint32_t* __restrict__ bm_row,
                                    unsigned long long* __restrict__ bitmap, uint32_t ld) {
    const uint32_t r = blockIdx.x;
    const uint32_t b = bm_row[r];
    if (b == kNoBitmap) return;
    unsigned long long* w = bitmap + (uint64_t)b * ld;
    for (uint64_t i = list_off[r] + threadIdx.x; i < list_off[r + 1]; i += blockDim.x) {
        const uint32_t x = list[i];
        atomicOr(&w[x >> 6], 1ull << (x & 63));
    }
}

int upload_bitmaps(storm_hip_corpus* c, hipStream_t s) {
    // (Re)allocate the store for every row with a bitmap plus the zero row.
    if (c->d_bitmap) { SH_HIP(hipFree(c->d_bitmap)); c->d_bitmap = nullptr; }
    const size_t bytes = ((size_t)c->n_bitmap_rows + 1) * c->ld * 8;
    SH_HIP(hipMalloc(&c->d_bitmap, bytes));
    c->bm_capacity = c->n_bitmap_rows;
    SH_HIP(hipMemsetAsync(c->d_bitmap, 0, bytes, s));
    SH_HIP(hipMemcpyAsync(c->d_bm_row, c->h_bm_row.data(), (size_t)c->n_rows * 4,
                          hipMemcpyHostToDevice, s));
    if (c->n_rows && c->has_lists) {
        hipLaunchKernelGGL(scatter_bits_kernel, dim3(c->n_rows), dim3(256), 0, s, c->d_list_off,
                           c->d_list, c->d_bm_row, (unsigned long long*)c->d_bitmap, c->ld);
        SH_HIP(hipGetLastError());
    } else if (c->n_rows && c->has_runs) {
        SH_TRY(fill_bitmaps_from_runs(c, s));
    } else if (c->n_rows && c->n_bitmap_rows) {
        set_error("bitmap rows requested for a corpus held without lists or runs");
        return STORM_HIP_ERR_INVALID;
    }
    return STORM_HIP_OK;
}

// Bit-sliced bitmaps: a workgroup per (slab, 8 words), its wave v on groups
// 8q + v and 8q + 4 + v. Lane k holds 8 words of sorted row 64g + k; for each
// of the 512 positions a ballot gathers the 64 rows' bits and lane t keeps the
// mask of position 64w + t. The waves fill the eight masks of the same
// positions (slab layout: position-major, kSlabGroups masks per position).
__global__ __launch_bounds__(256) void slice_bits_kernel(const uint32_t* __restrict__ perm,
                                                         const uint32_t* __restrict__ bm_row,
                                                         const uint64_t* __restrict__ bitmap,
                                                         const uint32_t* __restrict__ slice_of,
                                                         uint64_t* __restrict__ slice, uint32_t ld,
                                                         uint32_t groups) {
    const uint32_t lane = threadIdx.x & 63, v = threadIdx.x >> 6;
    const uint32_t w0 = blockIdx.x * 8;
    if (w0 >= ld) return;
    for (uint32_t q = blockIdx.y; kSlabGroups * q < groups; q += gridDim.y) {
        for (uint32_t h = 0; h < kSlabGroups; h += 4) {
            const uint32_t g = kSlabGroups * q + h + v;
            const uint32_t sl = g < groups ? slice_of[g] : kNoBitmap;
            if (sl == kNoBitmap) continue;
            // Padding positions of the density order map to the zero row.
            const uint64_t* src = bitmap + (uint64_t)bm_row[perm[64 * g + lane]] * ld + w0;
            uint64_t x[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) x[k] = src[k];
            uint64_t* dst = slice + ((uint64_t)(sl / kSlabGroups) * ld * 64 + (uint64_t)w0 * 64) *
                                        kSlabGroups + (sl % kSlabGroups);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                uint64_t mine = 0;
                for (uint32_t t = 0; t < 64; ++t) {
                    const uint64_t m = __ballot((x[k] >> t) & 1);
                    if (lane == t) mine = m;
                }
                dst[((uint64_t)k * 64 + lane) * kSlabGroups] = mine;
            }
        }
    }
}

// (Re)builds the slices of every 64-row group of the density order whose rows
// all hold bitmaps (rows past n are the zero row); groups 8q..8q+7 share slab
// q when any of them has slices.
int build_slices(storm_hip_corpus* c, hipStream_t s) {
    if (c->d_slice) { SH_HIP(hipFree(c->d_slice)); c->d_slice = nullptr; }
    if (c->d_slice_of) { SH_HIP(hipFree(c->d_slice_of)); c->d_slice_of = nullptr; }
    const uint64_t slab_bytes = (uint64_t)c->ld * 64 * kSlabGroups * 8;
    c->device_bytes -= (uint64_t)c->n_slices * slab_bytes;
    c->n_slices = 0;
    if (!c->want_slices || c->n_rows == 0) return STORM_HIP_OK;
    const uint32_t groups = (c->n_rows + 63) / 64;
    std::vector<uint32_t> of(groups, kNoBitmap);
    uint32_t ns = 0;
    for (uint32_t q = 0; kSlabGroups * q < groups; ++q) {
        bool any = false;
        for (uint32_t g = kSlabGroups * q; g < std::min(groups, kSlabGroups * (q + 1)); ++g) {
            bool all = true;
            for (uint32_t k = 64 * g; k < std::min(c->n_rows, 64 * g + 64) && all; ++k)
                all = c->h_bm_row[c->h_perm[k]] != kNoBitmap;
            if (all) { of[g] = kSlabGroups * ns + g % kSlabGroups; any = true; }
        }
        if (any) ++ns;
    }
    SH_TRY(upload_vec(c->d_slice_of, of, s, c->device_bytes));
    if (!ns) return STORM_HIP_OK;
    SH_HIP(hipMalloc(&c->d_slice, (size_t)ns * slab_bytes));
    SH_HIP(hipMemsetAsync(c->d_slice, 0, (size_t)ns * slab_bytes, s));
    c->n_slices = ns;
    c->device_bytes += (uint64_t)ns * slab_bytes;
    const dim3 grid(c->ld / 8, std::min<uint32_t>((groups + kSlabGroups - 1) / kSlabGroups, 65535));
    hipLaunchKernelGGL(slice_bits_kernel, grid, dim3(256), 0, s, c->d_perm, c->d_bm_row,
                       c->d_bitmap, c->d_slice_of, c->d_slice, c->ld, groups);
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

// Coarse zone map of one row per workgroup (STORM_HIP_BUILD_ZONES): bucket
// p >> shift of every position p the row holds, gathered in LDS from its list,
// else its runs, else its bitmap, then written out (zld words).
constexpr uint32_t kZoneBuckets = 16384, kZoneMaxWords = kZoneBuckets / 64 + 1;

__global__ __launch_bounds__(256) void zone_build_kernel(
    const uint64_t* __restrict__ list_off, const uint32_t* __restrict__ list,
    const uint64_t* __restrict__ run_off, const uint32_t* __restrict__ run_start,
    const uint32_t* __restrict__ run_end, const uint32_t* __restrict__ bm_row,
    const uint64_t* __restrict__ bitmap, uint32_t ld, uint32_t nw, uint32_t shift,
    uint32_t zld, uint64_t* __restrict__ zone) {
    __shared__ unsigned long long z[kZoneMaxWords];
    const uint32_t r = blockIdx.x, tid = threadIdx.x;
    for (uint32_t w = tid; w < zld; w += blockDim.x) z[w] = 0;
    __syncthreads();
    auto mark = [&](uint32_t b) { atomicOr(&z[b >> 6], 1ull << (b & 63)); };
    if (list) {
        for (uint64_t k = list_off[r] + tid; k < list_off[r + 1]; k += blockDim.x)
            mark(list[k] >> shift);
    } else if (run_off) {
        // A run [s, e) covers buckets s >> shift .. (e - 1) >> shift.
        for (uint64_t k = run_off[r] + tid; k < run_off[r + 1]; k += blockDim.x)
            for (uint32_t b = run_start[k] >> shift; b <= (run_end[k] - 1) >> shift; ++b) mark(b);
    } else if (bm_row[r] != kNoBitmap) {
        const uint64_t* bm = bitmap + (uint64_t)bm_row[r] * ld;
        for (uint32_t w = tid; w < nw; w += blockDim.x) {
            uint64_t v = bm[w];
            if (!v) continue;
            if (shift >= 6) {
                mark((w * 64u) >> shift);
            } else {
                while (v) {
                    mark((w * 64u + (uint32_t)__ffsll((unsigned long long)v) - 1) >> shift);
                    v &= v - 1;
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t w = tid; w < zld; w += blockDim.x) zone[(uint64_t)r * zld + w] = z[w];
}

// (Re)builds every row's zone map when the corpus asks for them.
int build_zones(storm_hip_corpus* c, hipStream_t s) {
    if (c->d_zone) {
        SH_HIP(hipFree(c->d_zone));
        c->d_zone = nullptr;
        c->device_bytes -= (uint64_t)c->n_rows * c->zone_ld * 8;
    }
    if (!c->want_zones || c->n_rows == 0 || c->dense) return STORM_HIP_OK;
    if (!c->d_list && !c->d_run_off && !c->d_bitmap) {
        set_error("zone maps: the rows hold no list, runs or bitmap");
        return STORM_HIP_ERR_INVALID;
    }
    // bench/bench_tile.cpp Occ::init: width = ceil(universe / buckets), shift
    // the smallest with 2^shift >= width, ((universe >> shift) + 1 + 63) / 64 words.
    const uint64_t width = ((uint64_t)c->universe + kZoneBuckets - 1) / kZoneBuckets;
    uint32_t shift = 0;
    while ((1ull << shift) < width) ++shift;
    c->zone_shift = shift;
    c->zone_ld = (uint32_t)((((uint64_t)c->universe >> shift) + 1 + 63) / 64);
    const uint64_t bytes = (uint64_t)c->n_rows * c->zone_ld * 8;
    SH_HIP(hipMalloc(&c->d_zone, bytes));
    c->device_bytes += bytes;
    const bool lists = c->d_list != nullptr && c->has_lists;
    const bool runs = !lists && c->d_run_off != nullptr;
    hipLaunchKernelGGL(zone_build_kernel, dim3(c->n_rows), dim3(256), 0, s,
                       lists ? c->d_list_off : nullptr, lists ? c->d_list : nullptr,
                       runs ? c->d_run_off : nullptr, c->d_run_start, c->d_run_end,
                       c->d_bm_row, c->d_bitmap, c->ld, c->n_words, shift, c->zone_ld,
                       c->d_zone);
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

int upload_host_arena(storm_hip_ctx* ctx, const HostArena& A, uint32_t flags,
                      storm_hip_corpus** out) {
    storm_hip_corpus* c = corpus_new(ctx, A.n, A.universe, flags);
    if (!c) return STORM_HIP_ERR_NOMEM;
    *out = c;   // freed by the caller on failure
    hipStream_t s = ctx->stream;
    uint64_t& B = c->device_bytes;
    SH_TRY(upload_vec(c->d_list_off, A.list_off, s, B));
    SH_TRY(upload_vec(c->d_list, A.list, s, B));
    SH_TRY(upload_vec(c->d_run_off, A.run_off, s, B));
    SH_TRY(upload_vec(c->d_run_start, A.run_start, s, B));
    SH_TRY(upload_vec(c->d_run_end, A.run_end, s, B));
    SH_TRY(upload_vec(c->d_ewah_off, A.ewah_off, s, B));
    SH_TRY(upload_vec(c->d_ewah, A.ewah, s, B));
    SH_TRY(upload_vec(c->d_mk_off, A.mk_off, s, B));
    SH_TRY(upload_vec(c->d_mk_word, A.mk_word, s, B));
    SH_TRY(upload_vec(c->d_mk_pos, A.mk_pos, s, B));
    SH_TRY(upload_vec(c->d_comp_off, A.comp_off, s, B));
    SH_TRY(upload_vec(c->d_comp, A.comp, s, B));
    SH_TRY(upload_vec(c->d_meta, A.meta, s, B));
    c->h_meta = A.meta;
    return corpus_finish(ctx, c, A.want_bitmap);
}

}  // namespace

storm_hip_corpus* corpus_new(storm_hip_ctx* ctx, uint32_t n, uint32_t universe, uint32_t flags) {
    storm_hip_corpus* c = new (std::nothrow) storm_hip_corpus();
    if (!c) return nullptr;
    c->ctx = ctx;
    c->dense = false;
    c->n_rows = n;
    c->universe = universe;
    c->n_words = (universe + 63) / 64;
    c->ld = (std::max(c->n_words, 1u) + kWordPad - 1) / kWordPad * kWordPad;
    c->has_runs = flags & STORM_HIP_BUILD_RUNS;
    c->has_ewah = flags & STORM_HIP_BUILD_EWAH;
    c->has_comp = flags & STORM_HIP_BUILD_COMP;
    c->want_slices = flags & STORM_HIP_BUILD_SLICES;
    c->want_zones = flags & STORM_HIP_BUILD_ZONES;
    return c;
}

std::vector<uint8_t> bitmap_rows(const std::vector<uint32_t>& meta, uint32_t n, uint32_t universe,
                                 uint32_t policy) {
    std::vector<uint8_t> want(n, 0);
    for (uint32_t r = 0; r < n; ++r)
        want[r] = policy == STORM_HIP_BITMAP_ALL ||
                  (policy == STORM_HIP_BITMAP_DENSE &&
                   (uint64_t)meta[(size_t)r * 8] * 64 >= (uint64_t)universe);
    return want;
}

namespace {

// Density order (stable by cardinality) and the identity order.
int upload_orders(storm_hip_corpus* c, hipStream_t s) {
    const uint32_t n = c->n_rows;
    uint64_t& B = c->device_bytes;
    const uint32_t npad = (n + kRowPad - 1) / kRowPad * kRowPad + kRowPad;
    c->h_perm.resize(n);
    std::iota(c->h_perm.begin(), c->h_perm.end(), 0u);
    std::stable_sort(c->h_perm.begin(), c->h_perm.end(), [&](uint32_t a, uint32_t b) {
        return c->h_meta[(size_t)a * 8] < c->h_meta[(size_t)b * 8];
    });
    std::vector<uint32_t> perm(npad, n), ident(npad, n);
    std::copy(c->h_perm.begin(), c->h_perm.end(), perm.begin());
    std::iota(ident.begin(), ident.begin() + n, 0u);
    // Padded like the orders: the sentinel row n maps to position n.
    std::vector<uint32_t> inv(npad, n);
    for (uint32_t k = 0; k < n; ++k) inv[c->h_perm[k]] = k;
    SH_TRY(upload_vec(c->d_perm, perm, s, B));
    SH_TRY(upload_vec(c->d_ident, ident, s, B));
    SH_TRY(upload_vec(c->d_inv, inv, s, B));
    // Cardinality, first and last position of the row at each position of
    // the density order (padding: empty), for the span tests of kernels that
    // walk positions.
    std::vector<uint32_t> span((size_t)npad * 4, 0);
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t* m = &c->h_meta[(size_t)c->h_perm[k] * 8];
        span[(size_t)k * 4] = m[0];
        span[(size_t)k * 4 + 1] = m[4];
        span[(size_t)k * 4 + 2] = m[5];
    }
    SH_TRY(upload_vec(c->d_pspan, span, s, B));
    return STORM_HIP_OK;
}

}  // namespace

// A store filled with slot = row (rows generated or copied in row order) is
// rearranged in place into density order (slot = position, as corpus_finish
// lays stores out), following the permutation's cycles with one spare row.
static int store_to_density_order(storm_hip_corpus* c, hipStream_t s) {
    const uint32_t n = c->n_rows;
    if (n < 2 || !c->d_bitmap || c->h_bm_row.size() < n) return STORM_HIP_OK;
    // Only a store still in row order (a store already laid out is left as it is).
    for (uint32_t r = 0; r < n; ++r)
        if (c->h_bm_row[r] != r) return STORM_HIP_OK;
    const size_t row_bytes = (size_t)c->ld * 8;
    uint64_t* spare = nullptr;
    SH_HIP(hipMalloc((void**)&spare, row_bytes));
    uint64_t* X = c->d_bitmap;
    auto slot = [&](uint32_t k) { return X + (uint64_t)k * c->ld; };
    std::vector<uint8_t> done(n, 0);
    hipError_t e = hipSuccess;
    // Slot k receives row perm[k], which sits in slot perm[k] until moved.
    for (uint32_t k = 0; k < n && e == hipSuccess; ++k) {
        if (done[k]) continue;
        done[k] = 1;
        if (c->h_perm[k] == k) continue;
        e = hipMemcpyAsync(spare, slot(k), row_bytes, hipMemcpyDeviceToDevice, s);
        uint32_t cur = k;
        while (e == hipSuccess) {
            const uint32_t src = c->h_perm[cur];
            if (src == k) {
                e = hipMemcpyAsync(slot(cur), spare, row_bytes, hipMemcpyDeviceToDevice, s);
                break;
            }
            e = hipMemcpyAsync(slot(cur), slot(src), row_bytes, hipMemcpyDeviceToDevice, s);
            done[src] = 1;
            cur = src;
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(spare);
    if (e != hipSuccess) {
        set_error("bitmap store reorder: %s", hipGetErrorString(e));
        return STORM_HIP_ERR_DEVICE;
    }
    for (uint32_t k = 0; k < n; ++k) c->h_bm_row[c->h_perm[k]] = k;
    SH_HIP(hipMemcpyAsync(c->d_bm_row, c->h_bm_row.data(), (size_t)n * 4, hipMemcpyHostToDevice,
                          s));
    return STORM_HIP_OK;
}

int corpus_finish_store(storm_hip_ctx* ctx, storm_hip_corpus* c) {
    hipStream_t s = ctx->stream;
    SH_TRY(upload_orders(c, s));
    SH_TRY(store_to_density_order(c, s));
    SH_TRY(build_indexes(c, s));
    SH_TRY(build_slices(c, s));
    SH_TRY(build_zones(c, s));
    SH_HIP(hipStreamSynchronize(s));
    return STORM_HIP_OK;
}

int corpus_finish(storm_hip_ctx* ctx, storm_hip_corpus* c, const std::vector<uint8_t>& want_bitmap) {
    const uint32_t n = c->n_rows;
    hipStream_t s = ctx->stream;
    uint64_t& B = c->device_bytes;
    // Density order: stable by cardinality (kernels/storm_allpairs.cpp:316-323),
    // padding positions map to the sentinel row id n (never read as data).
    SH_TRY(upload_orders(c, s));
    // Bitmap rows; the sentinel row id n maps to the zero row.
    c->h_bm_row.assign((size_t)n + 1, kNoBitmap);
    uint32_t nb = 0;
    // The store is laid out in density order: the rows of a tile of that
    // order (which the B x B kernels read as 256-row panels) are one
    // contiguous run of the store, not rows scattered over all of it.
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t r = c->h_perm[k];
        if (want_bitmap[r]) c->h_bm_row[r] = nb++;
    }
    c->n_bitmap_rows = nb;
    c->h_bm_row[n] = nb;   // zero row
    c->bitmaps_complete = (nb == n);
    SH_HIP(hipMalloc(&c->d_bm_row, ((size_t)n + 1) * 4));
    B += ((size_t)n + 1) * 4;
    SH_TRY(upload_bitmaps(c, s));
    SH_HIP(hipMemcpyAsync(c->d_bm_row + n, &c->h_bm_row[n], 4, hipMemcpyHostToDevice, s));
    B += ((size_t)nb + 1) * c->ld * 8;
    SH_TRY(build_indexes(c, s));
    SH_TRY(build_slices(c, s));
    SH_TRY(build_zones(c, s));
    SH_HIP(hipStreamSynchronize(s));
    return STORM_HIP_OK;
}

int general_materialize(storm_hip_ctx* ctx, storm_hip_corpus* c) {
    if (c->dense || c->bitmaps_complete) return STORM_HIP_OK;
    if (!c->has_lists && !c->has_runs) {
        set_error("materialise: rows without a bitmap hold no list or runs");
        return STORM_HIP_ERR_INVALID;
    }
    c->device_bytes -= ((size_t)c->n_bitmap_rows + 1) * c->ld * 8;
    // Every row, the store in density order (slot = position).
    for (uint32_t k = 0; k < c->n_rows; ++k) c->h_bm_row[c->h_perm[k]] = k;
    c->n_bitmap_rows = c->n_rows;
    c->h_bm_row[c->n_rows] = c->n_rows;
    SH_TRY(upload_bitmaps(c, ctx->stream));
    SH_HIP(hipMemcpyAsync(c->d_bm_row + c->n_rows, &c->h_bm_row[c->n_rows], 4,
                          hipMemcpyHostToDevice, ctx->stream));
    c->device_bytes += ((size_t)c->n_bitmap_rows + 1) * c->ld * 8;
    c->bitmaps_complete = true;
    SH_TRY(build_indexes(c, ctx->stream));
    SH_TRY(build_slices(c, ctx->stream));
    SH_HIP(hipStreamSynchronize(ctx->stream));
    return STORM_HIP_OK;
}

void general_free_arenas(storm_hip_corpus* c) {
    void* ptrs[] = {c->d_list_off, c->d_list, c->d_run_off, c->d_run_start, c->d_run_end,
                    c->d_ewah_off, c->d_ewah, c->d_mk_off, c->d_mk_word, c->d_mk_pos,
                    c->d_comp_off, c->d_comp, c->d_meta, c->d_bm_row, c->d_bitmap,
                    c->d_perm, c->d_ident, c->d_inv, c->d_pspan, c->d_occ,
                    c->d_slice, c->d_slice_of, c->d_rank, c->d_rank_row, c->d_occ_row,
                    c->d_zone};
    for (void* p : ptrs) if (p) (void)hipFree(p);
}

}  // namespace stormhip

using namespace stormhip;

extern "C" {

int storm_hip_corpus_from_lists(storm_hip_ctx* ctx, const uint64_t* offs, const uint32_t* pos,
                                uint32_t n_rows, uint32_t universe, uint32_t flags,
                                storm_hip_corpus** out) {
    if (!ctx || !out) return STORM_HIP_ERR_INVALID;
    *out = nullptr;
    if (!offs || (!pos && n_rows && offs[n_rows] > offs[0])) return STORM_HIP_ERR_NULL_VALUES;
    if (universe == 0 || (flags & 3u) == 3u) return STORM_HIP_ERR_INVALID;
    for (uint32_t r = 0; r < n_rows; ++r)
        if (offs[r + 1] < offs[r]) return STORM_HIP_ERR_INVALID;
    SH_HIP(hipSetDevice(ctx->device));
    if (flags & STORM_HIP_BUILD_ON_HOST) {
        HostArena A;
        SH_TRY(build_host_arena(offs, pos, n_rows, universe, flags, A));
        const int s = upload_host_arena(ctx, A, flags, out);
        if (s != STORM_HIP_OK && *out) { storm_hip_corpus_free(*out); *out = nullptr; }
        return s;
    }
    // Lists go up as they are (rebased to offset 0); everything else is built
    // on the device.
    storm_hip_corpus* c = corpus_new(ctx, n_rows, universe, flags);
    if (!c) return STORM_HIP_ERR_NOMEM;
    hipStream_t s = ctx->stream;
    std::vector<uint64_t> rebased(offs, offs + (size_t)n_rows + 1);
    for (auto& x : rebased) x -= offs[0];
    const uint64_t total = rebased[n_rows];
    int st = upload_vec(c->d_list_off, rebased, s, c->device_bytes);
    if (st == STORM_HIP_OK) {
        if (hipMalloc((void**)&c->d_list, std::max<uint64_t>(total, 1) * 4) != hipSuccess) {
            st = STORM_HIP_ERR_NOMEM;
        } else {
            c->device_bytes += std::max<uint64_t>(total, 1) * 4;
            if (total && hipMemcpyAsync(c->d_list, pos + offs[0], total * 4,
                                        hipMemcpyHostToDevice, s) != hipSuccess)
                st = STORM_HIP_ERR_DEVICE;
        }
    }
    if (st == STORM_HIP_OK) st = device_build(ctx, c, flags);
    if (st == STORM_HIP_OK)
        st = corpus_finish(ctx, c, bitmap_rows(c->h_meta, n_rows, universe, flags & 3u));
    if (st != STORM_HIP_OK) {
        storm_hip_corpus_free(c);
        return st;
    }
    *out = c;
    return STORM_HIP_OK;
}

int storm_hip_corpus_materialize_bitmaps(storm_hip_ctx* ctx, storm_hip_corpus* c) {
    if (!ctx || !c || c->ctx != ctx) return STORM_HIP_ERR_INVALID;
    SH_HIP(hipSetDevice(ctx->device));
    return general_materialize(ctx, c);
}

int storm_hip_corpus_arena_sizes(storm_hip_ctx* ctx, const storm_hip_corpus* c,
                                 uint64_t bytes[STORM_HIP_ARENA_COUNT]) {
    if (!ctx || !c || !bytes || c->ctx != ctx) return STORM_HIP_ERR_INVALID;
    if (c->dense) return STORM_HIP_ERR_UNSUPPORTED;
    SH_HIP(hipSetDevice(ctx->device));
    SH_HIP(hipStreamSynchronize(ctx->stream));
    const uint64_t n = c->n_rows;
    const uint64_t* offs[5] = {c->d_list_off, c->d_run_off, c->d_ewah_off, c->d_mk_off,
                               c->d_comp_off};
    uint64_t tot[5];
    for (int k = 0; k < 5; ++k) SH_HIP(hipMemcpy(&tot[k], offs[k] + n, 8, hipMemcpyDeviceToHost));
    bytes[STORM_HIP_ARENA_LIST_OFF] = (n + 1) * 8;
    bytes[STORM_HIP_ARENA_LIST] = tot[0] * 4;
    bytes[STORM_HIP_ARENA_RUN_OFF] = (n + 1) * 8;
    bytes[STORM_HIP_ARENA_RUN_START] = tot[1] * 4;
    bytes[STORM_HIP_ARENA_RUN_END] = tot[1] * 4;
    bytes[STORM_HIP_ARENA_EWAH_OFF] = (n + 1) * 8;
    bytes[STORM_HIP_ARENA_EWAH] = tot[2] * 8;
    bytes[STORM_HIP_ARENA_MK_OFF] = (n + 1) * 8;
    bytes[STORM_HIP_ARENA_MK_WORD] = tot[3] * 4;
    bytes[STORM_HIP_ARENA_MK_POS] = tot[3] * 4;
    bytes[STORM_HIP_ARENA_COMP_OFF] = (n + 1) * 8;
    bytes[STORM_HIP_ARENA_COMP] = tot[4] * 4;
    bytes[STORM_HIP_ARENA_META] = n * 32;
    bytes[STORM_HIP_ARENA_BM_ROW] = n * 4;
    bytes[STORM_HIP_ARENA_BITMAP] = (uint64_t)c->n_bitmap_rows * c->ld * 8;
    bytes[STORM_HIP_ARENA_RANK_ROW] = n * 4;
    bytes[STORM_HIP_ARENA_RANK] = (uint64_t)c->n_rank_rows * c->rank_ld * 8;
    bytes[STORM_HIP_ARENA_OCC_ROW] = n * 4;
    bytes[STORM_HIP_ARENA_OCC] = (uint64_t)c->n_occ_rows * c->occ_ld * 8;
    return STORM_HIP_OK;
}

int storm_hip_corpus_download_arena(storm_hip_ctx* ctx, const storm_hip_corpus* c, uint32_t which,
                                    void* dst, uint64_t bytes) {
    if (which >= STORM_HIP_ARENA_COUNT) return STORM_HIP_ERR_INVALID;
    uint64_t sizes[STORM_HIP_ARENA_COUNT];
    SH_TRY(storm_hip_corpus_arena_sizes(ctx, c, sizes));
    if (bytes != sizes[which]) {
        set_error("arena %u holds %llu bytes, %llu requested", which,
                  (unsigned long long)sizes[which], (unsigned long long)bytes);
        return STORM_HIP_ERR_INVALID;
    }
    if (bytes == 0) return STORM_HIP_OK;
    if (!dst) return STORM_HIP_ERR_NULL_VALUES;
    const void* src[STORM_HIP_ARENA_COUNT] = {
        c->d_list_off, c->d_list,  c->d_run_off, c->d_run_start, c->d_run_end,
        c->d_ewah_off, c->d_ewah,  c->d_mk_off,  c->d_mk_word,   c->d_mk_pos,
        c->d_comp_off, c->d_comp,  c->d_meta,    c->d_bm_row,    c->d_bitmap,
        c->d_rank_row, c->d_rank,  c->d_occ_row, c->d_occ};
    SH_HIP(hipMemcpy(dst, src[which], bytes, hipMemcpyDeviceToHost));
    return STORM_HIP_OK;
}

}  // extern "C"
