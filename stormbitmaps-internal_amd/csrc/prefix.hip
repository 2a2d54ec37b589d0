// Thresholded all-pairs by prefix-filtered candidate generation
// (bench/bench_threshold.cpp:210-290, after Bayardo, Ma & Srikant, WWW 2007):
// with the elements in one global total order and prefix length
// p(X) = |X| - ceil(t|X|) + 1, every pair with Jaccard >= t shares an element
// of their prefixes, so indexing the prefixes generates every pair that can
// be reported; each candidate is then counted exactly.
//
// On the device, one pass each:
//   1. document frequencies df(e) from the lists;
//   2. each row's prefix, one wavefront per row: the order is (floor(log2
//      df), id) -- rarest first like the reference's (df, id), but selectable
//      from a 32-bucket histogram of the row and one ordered sweep of its
//      sorted list, no per-row sort; any consistent total order is exact;
//   3. postings (element -> density-order positions of the rows whose
//      prefix holds it), CSR by a scan;
//   4. the reference's gate: estimated prefix work (2 sum p + sum_e
//      L_e(L_e - 1)/2 x mean |X|) over exact work (pairs x mean |X|)
//      (bench_threshold.cpp:170-200), prefix filtering below a cut measured
//      on the device (kPrefixGate), else the exact thresholded engine;
//   5. candidates: every row walks the postings of its prefix elements and
//      keeps the earlier rows (density order) that pass the length filter
//      min/max >= t, emitted as (i, j) keys, then sorted and made unique;
//   6. verification: one lane per candidate counts |A n B| exactly (a
//      merge, or the shorter list bisected in the longer) and reports
//      the pair when |A n B| > 0 and |A n B| / (|A| + |B| - |A n B|) >= t,
//      the reference's jac_ok in double precision.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "general.h"
#include "internal.h"

namespace stormhip {
namespace {

// The gate's cut. The reference's 5.0 (bench/bench_threshold.cpp:198) weighs
// both sides in CPU list-merge work. On the device the exact engine resolves a
// pair for about 0.06 ns (joins and stripes over whole tiles) while a verified
// candidate costs about 1.4 ns (emission, sort, merge), measured on a sparse
// skewed corpus of 100,000 rows: the crossover sits near 0.04 of the exact
// scan's estimated work.
constexpr double kPrefixGate = 0.04;

__device__ inline uint32_t df_bucket(uint32_t df) { return 31u - (uint32_t)__clz(df); }

__device__ inline void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__global__ void pf_df_kernel(const uint32_t* __restrict__ list, uint64_t total,
                             uint32_t* __restrict__ df) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(&df[list[k]], 1u);
}

// One wavefront per row: the plen = poff[r+1] - poff[r] first elements of the
// row in (bucket, id) order go to pfx[poff[r]..], in id order, and count once
// in pcnt of their element.
__global__ __launch_bounds__(256) void pf_select_kernel(
    const uint64_t* __restrict__ list_off, const uint32_t* __restrict__ list,
    const uint32_t* __restrict__ df, const uint64_t* __restrict__ poff, uint32_t n,
    uint32_t* __restrict__ pfx, uint32_t* __restrict__ pcnt) {
    __shared__ uint32_t hist[4][32];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t r = blockIdx.x * 4 + wave;
    if (r >= n) return;
    const uint32_t plen = (uint32_t)(poff[r + 1] - poff[r]);
    if (plen == 0) return;
    uint32_t* h = hist[wave];
    if (lane < 32) h[lane] = 0;
    wave_lds_fence();
    const uint64_t a = list_off[r], b = list_off[r + 1];
    for (uint64_t k = a + lane; k < b; k += 64) atomicAdd(&h[df_bucket(df[list[k]])], 1u);
    wave_lds_fence();
    // The bucket the prefix ends in, and how many of its elements it takes.
    uint32_t cum = 0, B = 31, below = 0;
    for (uint32_t q = 0; q < 32; ++q) {
        const uint32_t c = h[q];
        if (cum + c >= plen) { B = q; below = cum; break; }
        cum += c;
    }
    const uint32_t take = plen - below;
    const uint64_t lt = (1ull << lane) - 1;
    uint32_t seen_b = 0, out = 0;
    uint32_t* dst = pfx + poff[r];
    for (uint64_t k0 = a; k0 < b; k0 += 64) {
        const uint64_t k = k0 + lane;
        const bool in = k < b;
        const uint32_t e = in ? list[k] : 0u;
        const uint32_t bk = in ? df_bucket(df[e]) : 32u;
        const uint64_t is_b = __ballot(bk == B);
        const uint32_t rank_b = seen_b + (uint32_t)__popcll(is_b & lt);
        const bool sel = bk < B || (bk == B && rank_b < take);
        const uint64_t sm = __ballot(sel);
        if (sel) {
            dst[out + (uint32_t)__popcll(sm & lt)] = e;
            atomicAdd(&pcnt[e], 1u);
        }
        out += (uint32_t)__popcll(sm);
        seen_b += (uint32_t)__popcll(is_b);
    }
}

// Postings: every prefix element of row r records r's density position.
__global__ __launch_bounds__(256) void pf_fill_kernel(const uint64_t* __restrict__ poff,
                                                      const uint32_t* __restrict__ pfx,
                                                      const uint32_t* __restrict__ post_off,
                                                      const uint32_t* __restrict__ inv, uint32_t n,
                                                      uint32_t* __restrict__ fill,
                                                      uint32_t* __restrict__ post) {
    const uint32_t lane = threadIdx.x & 63, r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;
    const uint32_t q = inv[r];
    for (uint64_t k = poff[r] + lane; k < poff[r + 1]; k += 64) {
        const uint32_t e = pfx[k];
        post[post_off[e] + atomicAdd(&fill[e], 1u)] = q;
    }
}

// sum_e L_e (L_e - 1) / 2 over the posting lengths (the gate's candidate term).
__global__ void pf_sumsq_kernel(const uint32_t* __restrict__ pcnt, uint32_t m,
                                double* __restrict__ out) {
    double v = 0.0;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        const double L = (double)pcnt[e];
        v += L * (L - 1.0) * 0.5;
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v != 0.0) atomicAdd(out, v);
}

// Candidates: one wavefront per row r walks the postings of r's prefix
// elements one element at a time, its lanes striding the element's list (a
// popular element's long list is shared by the wave, not left to one lane);
// an earlier row (smaller density position) passing the length filter yields
// key (min(r, j) << 32 | max(r, j)). The wave counts (out null) or writes at
// slots taken with one atomic per step.
__global__ __launch_bounds__(256) void pf_cand_kernel(
    const uint64_t* __restrict__ poff, const uint32_t* __restrict__ pfx,
    const uint32_t* __restrict__ post_off, const uint32_t* __restrict__ post,
    const uint32_t* __restrict__ inv, const uint32_t* __restrict__ perm,
    const uint32_t* __restrict__ meta, double t, uint32_t n, unsigned long long* __restrict__ count,
    uint64_t* __restrict__ out, uint64_t cap) {
    const uint32_t lane = threadIdx.x & 63, r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;
    const uint32_t q = inv[r];
    const double ci = (double)meta[(uint64_t)r * 8];
    const uint64_t lt = (1ull << lane) - 1;
    for (uint64_t k = poff[r]; k < poff[r + 1]; ++k) {
        const uint32_t e = pfx[k];
        const uint32_t lo = post_off[e], hi = post_off[e + 1];
        for (uint32_t x0 = lo; x0 < hi; x0 += 64) {
            const uint32_t x = x0 + lane;
            bool emit = false;
            uint32_t j = 0;
            if (x < hi) {
                const uint32_t qj = post[x];
                if (qj < q) {
                    j = perm[qj];
                    const double cj = (double)meta[(uint64_t)j * 8];
                    emit = fmin(ci, cj) / fmax(ci, cj) >= t;
                }
            }
            const uint64_t em = __ballot(emit);
            if (!em) continue;
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(count, (unsigned long long)__popcll(em));
            base = __shfl(base, 0, 64);
            if (emit && out) {
                const uint64_t slot = base + (uint64_t)__popcll(em & lt);
                if (slot < cap)
                    out[slot] = ((uint64_t)min(r, j) << 32) | (uint64_t)max(r, j);
            }
        }
    }
}

// |A n B| of two sorted lists by a merge, or by bisecting the shorter list's
// elements in the longer when that is cheaper (na log2 nb < na + nb).
__device__ inline uint32_t pf_intersect(const uint32_t* A, uint32_t na, const uint32_t* B,
                                        uint32_t nb) {
    if (na > nb) { const uint32_t* T = A; A = B; B = T; const uint32_t m = na; na = nb; nb = m; }
    uint32_t c = 0;
    if ((uint64_t)na * (32u - (uint32_t)__clz(nb | 1u)) < (uint64_t)na + nb) {
        for (uint32_t k = 0; k < na; ++k) {
            const uint32_t x = A[k];
            uint32_t lo = 0, hi = nb;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (B[mid] < x) lo = mid + 1; else hi = mid;
            }
            c += (lo < nb && B[lo] == x) ? 1u : 0u;
        }
        return c;
    }
    uint32_t x = 0, y = 0;
    while (x < na && y < nb) {
        const uint32_t a = A[x], b = B[y];
        c += a == b ? 1u : 0u;
        x += a <= b ? 1u : 0u;
        y += b <= a ? 1u : 0u;
    }
    return c;
}

// Verification: one lane per unique candidate (i, j), grid-stride over the
// candidates; the pairs that pass are appended with one atomic per wave.
__global__ __launch_bounds__(256) void pf_verify_kernel(
    const uint64_t* __restrict__ keys, const unsigned int* __restrict__ n_keys,
    const uint64_t* __restrict__ list_off, const uint32_t* __restrict__ list, double t,
    PairHit* __restrict__ hits, unsigned long long* __restrict__ n_hits,
    unsigned long long* __restrict__ sum) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1;
    const uint64_t nk = *n_keys;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    // Every lane of a wave runs the same number of iterations (the ballots).
    const uint64_t first = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
    for (uint64_t w0 = first; w0 < nk; w0 += stride) {
        const uint64_t w = w0 + lane;
        bool ok = false;
        uint32_t i = 0, j = 0, c = 0;
        if (w < nk) {
            const uint64_t key = keys[w];
            i = (uint32_t)(key >> 32);
            j = (uint32_t)key;
            const uint32_t na = (uint32_t)(list_off[i + 1] - list_off[i]);
            const uint32_t nb = (uint32_t)(list_off[j + 1] - list_off[j]);
            c = pf_intersect(list + list_off[i], na, list + list_off[j], nb);
            // bench/bench_threshold.cpp jac_ok: u = |A| + |B| - |A n B| in double.
            ok = c > 0 && (double)c / ((double)na + (double)nb - (double)c) >= t;
        }
        const uint64_t m = __ballot(ok);
        if (!m) continue;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(n_hits, (unsigned long long)__popcll(m));
        base = __shfl(base, 0, 64);
        const uint64_t part = wave_sum_u64(ok ? (uint64_t)c : 0ull);
        if (lane == 0) atomicAdd(sum, (unsigned long long)part);
        if (ok) hits[base + (uint64_t)__popcll(m & lt)] = PairHit{i, j, c};
    }
}

uint32_t grid_of(uint64_t threads, uint32_t block, uint32_t cap) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((threads + block - 1) / block, cap));
}

// Device buffers of one call, freed on every exit path.
struct Bufs {
    std::vector<void*> p;
    template <class T>
    int get(T** out, uint64_t count) {
        *out = nullptr;
        SH_HIP(hipMalloc((void**)out, std::max<uint64_t>(1, count) * sizeof(T)));
        p.push_back(*out);
        return STORM_HIP_OK;
    }
    ~Bufs() {
        for (void* x : p) (void)hipFree(x);
    }
};

int prefix_call(storm_hip_ctx* ctx, storm_hip_corpus* c, double t, uint32_t mode,
                const storm_hip_options* exact_opt, storm_hip_stats& s, uint64_t* n_pairs,
                storm_hip_prefix_info& info) {
    const uint32_t n = c->n_rows;
    hipStream_t st = ctx->stream;
    // The per-row kernels launch a wavefront per row (64 n work-items).
    if (n > (1u << 26)) {
        set_error("prefix filter: %u rows exceed the per-row launches", n);
        return STORM_HIP_ERR_RANGE;
    }
    // Prefix lengths (bench_threshold.cpp:232-236) and the gate's host terms.
    std::vector<uint64_t> poff(n + 1, 0);
    uint64_t total = 0;   // list elements (a list holds its row's cardinality)
    for (uint32_t r = 0; r < n; ++r) {
        const uint32_t ci = c->h_meta[(size_t)r * 8];
        const long long p = std::max<long long>(0, (long long)ci - (long long)std::ceil(t * ci) + 1);
        poff[r + 1] = poff[r] + std::min<uint64_t>((uint64_t)p, ci);
        total += ci;
    }
    const double mean_s = (double)total / n;
    const uint64_t sump = poff[n];
    const double pairs = (double)n * (n - 1) / 2.0;
    info.prefix_elements = sump;
    if (sump >= (1ull << 32)) {
        set_error("prefix filter: %llu prefix elements exceed the 32-bit postings",
                  (unsigned long long)sump);
        return STORM_HIP_ERR_RANGE;
    }
    const uint32_t U = c->universe;
    Bufs B;
    uint32_t *df = nullptr, *pcnt = nullptr, *pfx = nullptr, *post_off = nullptr,
             *fill = nullptr, *post = nullptr;
    uint64_t* d_poff = nullptr;
    double* d_sumsq = nullptr;
    SH_TRY(B.get(&df, U));
    SH_TRY(B.get(&pcnt, (uint64_t)U + 1));
    SH_TRY(B.get(&post_off, (uint64_t)U + 1));
    SH_TRY(B.get(&d_poff, (uint64_t)n + 1));
    SH_TRY(B.get(&pfx, sump));
    SH_TRY(B.get(&d_sumsq, 1));
    SH_HIP(hipMemsetAsync(df, 0, (size_t)U * 4, st));
    SH_HIP(hipMemsetAsync(pcnt, 0, ((size_t)U + 1) * 4, st));
    SH_HIP(hipMemsetAsync(d_sumsq, 0, 8, st));
    SH_HIP(hipMemcpyAsync(d_poff, poff.data(), poff.size() * 8, hipMemcpyHostToDevice, st));
    const auto t0 = std::chrono::steady_clock::now();
    SH_HIP(hipEventRecord(ctx->ev[0], st));
    hipLaunchKernelGGL(pf_df_kernel, dim3(grid_of(total, 256, 16384)), dim3(256), 0, st,
                       c->d_list, total, df);
    hipLaunchKernelGGL(pf_select_kernel, dim3((n + 3) / 4), dim3(256), 0, st, c->d_list_off,
                       c->d_list, df, d_poff, n, pfx, pcnt);
    hipLaunchKernelGGL(pf_sumsq_kernel, dim3(grid_of(U, 256, 4096)), dim3(256), 0, st, pcnt, U,
                       d_sumsq);
    SH_HIP(hipGetLastError());
    double sumsq = 0.0;
    SH_HIP(hipMemcpyAsync(&sumsq, d_sumsq, 8, hipMemcpyDeviceToHost, st));
    SH_HIP(hipStreamSynchronize(st));
    info.est_ratio = pairs > 0 && mean_s > 0 ? (2.0 * (double)sump + sumsq * mean_s) /
                                                   (pairs * mean_s)
                                             : 1e9;
    info.used_prefix = mode == STORM_HIP_PREFIX_ALWAYS ||
                       (mode == STORM_HIP_PREFIX_GATED && info.est_ratio < kPrefixGate);
    if (!info.used_prefix) return STORM_HIP_ERR_UNSUPPORTED;   // the caller runs the exact engine
    // Postings: CSR offsets by an exclusive scan of the counts (pcnt[U] = 0).
    size_t tmp_bytes = 0;
    SH_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, pcnt, post_off, U + 1, st));
    uint8_t* tmp = nullptr;
    SH_TRY(B.get(&tmp, tmp_bytes));
    SH_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, pcnt, post_off, U + 1, st));
    SH_TRY(B.get(&fill, U));
    SH_TRY(B.get(&post, sump));
    SH_HIP(hipMemsetAsync(fill, 0, (size_t)U * 4, st));
    hipLaunchKernelGGL(pf_fill_kernel, dim3((n + 3) / 4), dim3(256), 0, st, d_poff, pfx, post_off,
                       c->d_inv, n, fill, post);
    // Candidates: count, then emit into a buffer of exactly that size.
    unsigned long long* d_cnt = nullptr;
    SH_TRY(B.get(&d_cnt, 2));
    SH_HIP(hipMemsetAsync(d_cnt, 0, 16, st));
    hipLaunchKernelGGL(pf_cand_kernel, dim3((n + 3) / 4), dim3(256), 0, st, d_poff, pfx, post_off,
                       post, c->d_inv, c->d_perm, c->d_meta, t, n, d_cnt, (uint64_t*)nullptr,
                       (uint64_t)0);
    SH_HIP(hipGetLastError());
    unsigned long long ncand = 0;
    SH_HIP(hipMemcpyAsync(&ncand, d_cnt, 8, hipMemcpyDeviceToHost, st));
    SH_HIP(hipStreamSynchronize(st));
    info.candidates = ncand;
    if (ncand >= (1ull << 31)) {
        set_error("prefix filter: %llu candidates exceed one sort", ncand);
        return STORM_HIP_ERR_RANGE;
    }
    const int nc = (int)ncand;
    uint64_t *keys = nullptr, *keys2 = nullptr;
    unsigned int* n_unique = nullptr;
    SH_TRY(B.get(&keys, ncand));
    SH_TRY(B.get(&keys2, ncand));
    SH_TRY(B.get(&n_unique, 1));
    SH_HIP(hipMemsetAsync(d_cnt, 0, 8, st));
    SH_HIP(hipMemsetAsync(n_unique, 0, 4, st));
    hipLaunchKernelGGL(pf_cand_kernel, dim3((n + 3) / 4), dim3(256), 0, st, d_poff, pfx, post_off,
                       post, c->d_inv, c->d_perm, c->d_meta, t, n, d_cnt, keys, ncand);
    SH_HIP(hipGetLastError());
    if (nc > 0) {
        size_t sb = 0, ub = 0;
        SH_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, sb, keys, keys2, nc, 0, 64, st));
        SH_HIP(hipcub::DeviceSelect::Unique(nullptr, ub, keys2, keys, n_unique, nc, st));
        uint8_t* tmp2 = nullptr;
        SH_TRY(B.get(&tmp2, std::max(sb, ub)));
        SH_HIP(hipcub::DeviceRadixSort::SortKeys(tmp2, sb, keys, keys2, nc, 0, 64, st));
        SH_HIP(hipcub::DeviceSelect::Unique(tmp2, ub, keys2, keys, n_unique, nc, st));
    }
    unsigned int nu = 0;
    SH_HIP(hipMemcpyAsync(&nu, n_unique, 4, hipMemcpyDeviceToHost, st));
    SH_HIP(hipStreamSynchronize(st));
    info.unique_candidates = nu;
    // Verification into a pair buffer as large as the candidates.
    PairHit* hits = nullptr;
    unsigned long long* d_hits_sum = nullptr;
    SH_TRY(B.get(&hits, nu));
    SH_TRY(B.get(&d_hits_sum, 2));
    SH_HIP(hipMemsetAsync(d_hits_sum, 0, 16, st));
    if (nu)
        hipLaunchKernelGGL(pf_verify_kernel, dim3(grid_of(nu, 256, 65536)), dim3(256), 0, st, keys,
                           n_unique, c->d_list_off, c->d_list, t, hits, d_hits_sum,
                           d_hits_sum + 1);
    SH_HIP(hipGetLastError());
    SH_HIP(hipEventRecord(ctx->ev[1], st));
    unsigned long long hs[2] = {0, 0};
    SH_HIP(hipMemcpyAsync(hs, d_hits_sum, 16, hipMemcpyDeviceToHost, st));
    SH_HIP(hipStreamSynchronize(st));
    std::vector<PairHit> h(hs[0]);
    if (hs[0]) SH_HIP(hipMemcpy(h.data(), hits, hs[0] * sizeof(PairHit), hipMemcpyDeviceToHost));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    s.ms_device = ms;
    std::sort(h.begin(), h.end(), [](const PairHit& a, const PairHit& b) {
        return a.i != b.i ? a.i < b.i : a.j < b.j;
    });
    c->h_hits.resize(h.size() * 3);
    for (size_t k = 0; k < h.size(); ++k) {
        c->h_hits[3 * k] = h[k].i;
        c->h_hits[3 * k + 1] = h[k].j;
        c->h_hits[3 * k + 2] = h[k].count;
    }
    c->hits_valid = true;
    *n_pairs = h.size();
    s.sum = hs[1];
    s.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                     .count();
    (void)exact_opt;
    return STORM_HIP_OK;
}

}  // namespace
}  // namespace stormhip

using namespace stormhip;

extern "C" {

int storm_hip_allpairs_threshold_prefix(storm_hip_ctx* ctx, storm_hip_corpus* c, double t,
                                        uint32_t mode, const storm_hip_options* exact_opt,
                                        storm_hip_stats* st, uint64_t* n_pairs,
                                        storm_hip_prefix_info* info) {
    if (!ctx || !c || c->ctx != ctx || !n_pairs) return STORM_HIP_ERR_INVALID;
    *n_pairs = 0;
    if (!(t > 0.0 && t <= 1.0)) {
        set_error("Jaccard threshold must be in (0, 1] (got %g)", t);
        return STORM_HIP_ERR_INVALID;
    }
    if (mode > STORM_HIP_PREFIX_EXACT) return STORM_HIP_ERR_INVALID;
    storm_hip_prefix_info local_info{};
    storm_hip_prefix_info& inf = info ? *info : local_info;
    std::memset(&inf, 0, sizeof inf);
    storm_hip_stats local{};
    storm_hip_stats& s = st ? *st : local;
    std::memset(&s, 0, sizeof s);
    const bool lists = !c->dense && c->has_lists && c->d_list && c->n_rows >= 2;
    if (mode == STORM_HIP_PREFIX_ALWAYS && !lists) {
        set_error("prefix filtering reads the rows' lists: this corpus holds none");
        return STORM_HIP_ERR_UNSUPPORTED;
    }
    if (mode != STORM_HIP_PREFIX_EXACT && lists) {
        SH_HIP(hipSetDevice(ctx->device));
        c->hits_valid = false;
        c->h_hits.clear();
        const int rc = prefix_call(ctx, c, t, mode, exact_opt, s, n_pairs, inf);
        s.pairs = (uint64_t)c->n_rows * (c->n_rows - 1) / 2;
        // Gated calls whose candidates outgrow one sort run the exact engine.
        const bool too_many = rc == STORM_HIP_ERR_RANGE && mode == STORM_HIP_PREFIX_GATED;
        if (!too_many && (rc != STORM_HIP_ERR_UNSUPPORTED || inf.used_prefix)) return rc;
    }
    // The gate chose the exact scan (or it was asked for): the thresholded engine.
    storm_hip_options o{};
    if (exact_opt) o = *exact_opt;
    else { o.policy = STORM_HIP_POLICY_ANALYTIC; o.tile = 256; }
    const double est = inf.est_ratio;
    const uint64_t pe = inf.prefix_elements;
    const int rc = storm_hip_allpairs_threshold(ctx, c, &o, t, &s, n_pairs);
    inf.used_prefix = 0;
    inf.est_ratio = est;
    inf.prefix_elements = pe;
    return rc;
}

}  // extern "C"
