// C ABI entry points: contexts, dense corpora, the all-pairs call for dense
// corpora, result access. General (mixed-representation) corpora are handled
// by general.hip through the same storm_hip_allpairs entry point.
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

#include "device_common.h"
#include "internal.h"

namespace stormhip {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int general_allpairs(storm_hip_ctx* ctx, storm_hip_corpus* c, const storm_hip_options* o,
                     storm_hip_stats* st, unsigned long long* d_sum);
void general_free(storm_hip_corpus* c);
void default_device_model(DeviceCostModel& m);
void selector_destroy(storm_hip_ctx* ctx);

// Bernoulli(p) bitmap generator: one thread per 64-bit word. Draw k of
// Rng(seed) is mix(seed + (k+1)*golden); bit b of row r is draw r*n_bits+b.
__global__ void gen_bernoulli_kernel(uint64_t* __restrict__ X, uint32_t ld, uint32_t n_rows,
                                     uint32_t n_rows_pad, uint32_t n_bits, double p53,
                                     uint64_t seed) {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = (uint64_t)n_rows_pad * ld;
    if (gid >= total) return;
    const uint32_t r = (uint32_t)(gid / ld), w = (uint32_t)(gid % ld);
    uint64_t word = 0;
    if (r < n_rows && (uint64_t)w * 64 < n_bits) {
        const uint32_t nb = min(64u, n_bits - w * 64);
        uint64_t s = seed + ((uint64_t)r * n_bits + (uint64_t)w * 64) * kGolden;
        for (uint32_t b = 0; b < nb; ++b) {
            s += kGolden;
            const uint64_t m = splitmix_mix(s) >> 11;
            if ((double)m < p53) word |= 1ull << b;
        }
    }
    X[gid] = word;
}

// Row sums of a packed upper-triangle result: one block per row.
__global__ void tri_row_sums_kernel(const uint32_t* __restrict__ tri, uint32_t n,
                                    const uint32_t* __restrict__ perm,
                                    unsigned long long* __restrict__ out) {
    const uint32_t i = blockIdx.x;
    uint64_t s = 0;
    const uint64_t base = tri_index(i, i + 1, n);
    for (uint32_t j = i + 1 + threadIdx.x; j < n; j += blockDim.x) s += tri[base + (j - i - 1)];
    for (uint32_t j = threadIdx.x; j < i; j += blockDim.x) s += tri[tri_index(j, i, n)];
    s = wave_sum_u64(s);
    __shared__ unsigned long long part[16];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) t += part[w];
        out[perm ? perm[i] : i] = t;   // a position-packed triangle: row perm[i]
    }
}

static uint32_t round_up(uint32_t x, uint32_t m) { return (x + m - 1) / m * m; }

static int alloc_dense(storm_hip_ctx* ctx, uint32_t n_rows, uint32_t n_words,
                       storm_hip_corpus** out) {
    storm_hip_corpus* c = new (std::nothrow) storm_hip_corpus();
    if (!c) return STORM_HIP_ERR_NOMEM;
    c->ctx = ctx;
    c->dense = true;
    c->n_rows = n_rows;
    c->n_words = n_words;
    c->universe = n_words * 64;
    c->ld = round_up(n_words ? n_words : 1, kWordPad);
    c->n_rows_pad = round_up(n_rows ? n_rows : 1, kRowPad);
    const size_t bytes = (size_t)c->n_rows_pad * c->ld * 8;
    hipError_t e = hipMalloc(&c->d_words, bytes);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu) for dense corpus: %s", bytes, hipGetErrorString(e));
        delete c;
        return e == hipErrorOutOfMemory ? STORM_HIP_ERR_NOMEM : STORM_HIP_ERR_DEVICE;
    }
    c->device_bytes = bytes;
    c->bitmaps_complete = true;
    *out = c;
    return STORM_HIP_OK;
}

}  // namespace stormhip

using namespace stormhip;

extern "C" {

const char* storm_hip_strerror(int s) {
    switch (s) {
        case STORM_HIP_OK: return "ok";
        case STORM_HIP_ERR_INVALID: return "invalid argument or handle";
        case STORM_HIP_ERR_NULL_VALUES: return "null data pointer";
        case STORM_HIP_ERR_DEVICE: return "HIP device error";
        case STORM_HIP_ERR_NOMEM: return "out of memory";
        case STORM_HIP_ERR_UNSUPPORTED: return "unsupported";
        case STORM_HIP_ERR_RANGE: return "size out of range";
    }
    return "unknown status";
}

const char* storm_hip_version(void) { return "storm_hip 0.1 gfx950 abi 3"; }

const char* storm_hip_last_error(void) { return g_err; }

int storm_hip_device_count(int* n) {
    if (!n) return STORM_HIP_ERR_INVALID;
    int k = 0;
    if (hipGetDeviceCount(&k) != hipSuccess) k = 0;
    *n = k;
    return STORM_HIP_OK;
}

int storm_hip_ctx_create(int device, uint32_t flags, storm_hip_ctx** out) {
    (void)flags;
    if (!out) return STORM_HIP_ERR_INVALID;
    *out = nullptr;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) {
        set_error("no HIP device visible");
        return STORM_HIP_ERR_DEVICE;
    }
    if (device < 0 || device >= nd) {
        set_error("device %d out of range (%d visible)", device, nd);
        return STORM_HIP_ERR_INVALID;
    }
    SH_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    SH_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error("device %d is %s; this build targets gfx950 only", device, prop.gcnArchName);
        return STORM_HIP_ERR_UNSUPPORTED;
    }
    SH_TRY(noalloc_register_device(device));
    storm_hip_ctx* c = new (std::nothrow) storm_hip_ctx();
    if (!c) return STORM_HIP_ERR_NOMEM;
    c->device = device;
    c->cu_count = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->aux[0], hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->aux[1], hipStreamNonBlocking) != hipSuccess) {
        for (hipStream_t s : {c->stream, c->aux[0], c->aux[1]})
            if (s) (void)hipStreamDestroy(s);
        delete c;
        set_error("hipStreamCreate failed");
        return STORM_HIP_ERR_DEVICE;
    }
    for (auto& e : c->ev) hipEventCreate(&e);
    hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming);
    for (auto& e : c->join_ev) hipEventCreateWithFlags(&e, hipEventDisableTiming);
    default_device_model(c->model);
    *out = c;
    return STORM_HIP_OK;
}

void storm_hip_ctx_free(storm_hip_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (hipStream_t s : ctx->aux) if (s) hipStreamSynchronize(s);
    selector_destroy(ctx);
    for (auto& e : ctx->ev) if (e) hipEventDestroy(e);
    if (ctx->fork_ev) hipEventDestroy(ctx->fork_ev);
    for (auto& e : ctx->join_ev) if (e) hipEventDestroy(e);
    for (hipStream_t s : ctx->aux) if (s) hipStreamDestroy(s);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

int storm_hip_sync(storm_hip_ctx* ctx) {
    if (!ctx) return STORM_HIP_ERR_INVALID;
    SH_HIP(hipStreamSynchronize(ctx->stream));
    return STORM_HIP_OK;
}

int storm_hip_corpus_upload_contig(storm_hip_ctx* ctx, const uint64_t* words, uint32_t n_rows,
                                   uint32_t n_words, storm_hip_corpus** out) {
    if (!ctx || !out) return STORM_HIP_ERR_INVALID;
    if (!words && n_rows && n_words) return STORM_HIP_ERR_NULL_VALUES;
    if (n_words > (1u << 26)) return STORM_HIP_ERR_RANGE;
    SH_HIP(hipSetDevice(ctx->device));
    storm_hip_corpus* c = nullptr;
    SH_TRY(alloc_dense(ctx, n_rows, n_words, &c));
    *out = c;
    SH_HIP(hipMemsetAsync(c->d_words, 0, c->device_bytes, ctx->stream));
    if (n_rows && n_words)
        SH_HIP(hipMemcpy2DAsync(c->d_words, (size_t)c->ld * 8, words, (size_t)n_words * 8,
                                (size_t)n_words * 8, n_rows, hipMemcpyHostToDevice, ctx->stream));
    SH_HIP(hipStreamSynchronize(ctx->stream));
    return STORM_HIP_OK;
}

int storm_hip_corpus_gen_bernoulli(storm_hip_ctx* ctx, uint32_t n_rows, uint32_t n_bits,
                                   double p, uint64_t seed, storm_hip_corpus** out) {
    if (!ctx || !out || !(p >= 0.0 && p <= 1.0)) return STORM_HIP_ERR_INVALID;
    if (n_bits > (1u << 31)) return STORM_HIP_ERR_RANGE;
    SH_HIP(hipSetDevice(ctx->device));
    storm_hip_corpus* c = nullptr;
    SH_TRY(alloc_dense(ctx, n_rows, (n_bits + 63) / 64, &c));
    c->universe = n_bits;
    *out = c;
    const uint64_t total = (uint64_t)c->n_rows_pad * c->ld;
    const uint64_t seed0 = seed ? seed : kGolden;
    hipLaunchKernelGGL(gen_bernoulli_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0,
                       ctx->stream, c->d_words, c->ld, n_rows, c->n_rows_pad, n_bits,
                       p * 9007199254740992.0, seed0);
    SH_HIP(hipGetLastError());
    SH_HIP(hipStreamSynchronize(ctx->stream));
    return STORM_HIP_OK;
}

int storm_hip_corpus_download_words(storm_hip_ctx* ctx, const storm_hip_corpus* c,
                                    uint32_t row0, uint32_t n_rows, uint64_t* out) {
    if (!ctx || !c || !c->dense) return STORM_HIP_ERR_INVALID;
    if (!out && n_rows) return STORM_HIP_ERR_NULL_VALUES;
    if ((uint64_t)row0 + n_rows > c->n_rows) return STORM_HIP_ERR_RANGE;
    if (!n_rows || !c->n_words) return STORM_HIP_OK;
    SH_HIP(hipSetDevice(ctx->device));
    SH_HIP(hipMemcpy2DAsync(out, (size_t)c->n_words * 8, c->d_words + (size_t)row0 * c->ld,
                            (size_t)c->ld * 8, (size_t)c->n_words * 8, n_rows,
                            hipMemcpyDeviceToHost, ctx->stream));
    SH_HIP(hipStreamSynchronize(ctx->stream));
    return STORM_HIP_OK;
}

void storm_hip_corpus_free(storm_hip_corpus* c) {
    if (!c) return;
    if (c->ctx) {
        hipSetDevice(c->ctx->device);
        hipStreamSynchronize(c->ctx->stream);
    }
    if (c->d_words) hipFree(c->d_words);
    if (c->d_tri) hipFree(c->d_tri);
    general_free(c);
    delete c;
}

uint32_t storm_hip_corpus_rows(const storm_hip_corpus* c) { return c ? c->n_rows : 0; }
uint64_t storm_hip_corpus_device_bytes(const storm_hip_corpus* c) {
    return c ? c->device_bytes + c->tri_count * 4 : 0;
}

int storm_hip_allpairs(storm_hip_ctx* ctx, storm_hip_corpus* c, const storm_hip_options* opt,
                       storm_hip_stats* st) {
    if (!ctx || !c || c->ctx != ctx) return STORM_HIP_ERR_INVALID;
    storm_hip_options o{};
    if (opt) o = *opt;
    if (o.policy > STORM_HIP_POLICY_PER_PAIR || o.out_mode > STORM_HIP_OUT_UPPER_TRI_DENSITY ||
        o.bb_kernel > STORM_HIP_BB_MFMA_FP4 || o.fixed_cell >= STORM_HIP_NCELL)
        return STORM_HIP_ERR_INVALID;
    storm_hip_stats local{};
    storm_hip_stats& s = st ? *st : local;
    std::memset(&s, 0, sizeof s);
    const auto t0 = std::chrono::steady_clock::now();
    SH_HIP(hipSetDevice(ctx->device));
    const uint64_t n = c->n_rows;
    s.pairs = n ? n * (n - 1) / 2 : 0;

    unsigned long long* d_sum = nullptr;
    SH_HIP(hipMallocAsync((void**)&d_sum, 8, ctx->stream));
    SH_HIP(hipMemsetAsync(d_sum, 0, 8, ctx->stream));
    uint32_t* tri = nullptr;
    if (o.out_mode != STORM_HIP_OUT_SUM) {
        if (c->tri_count != s.pairs) {
            if (c->d_tri) { SH_HIP(hipFree(c->d_tri)); c->d_tri = nullptr; c->tri_count = 0; }
            if (s.pairs) SH_HIP(hipMalloc(&c->d_tri, s.pairs * 4));
            c->tri_count = s.pairs;
        }
        tri = c->d_tri;
    }
    c->tri_valid = false;
    SH_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    int rc = STORM_HIP_OK;
    if (c->dense) {
        // Every row is a bitmap: the only cell is B x B. All policies agree
        // here; the span test would only skip work (AllBitmap never applies it).
        SH_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
        rc = launch_bb_dense(o.bb_kernel, ctx->mfma, ctx->stream, c->d_words, c->ld, c->n_words,
                             nullptr, c->n_rows, tri, d_sum, (uint32_t)ctx->cu_count);
        SH_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
        s.cell_pairs[STORM_HIP_CELL_BB] = s.pairs;
        s.bb_launches = 1;
    } else {
        rc = general_allpairs(ctx, c, &o, &s, d_sum);
    }
    SH_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    uint64_t h_sum = 0;
    SH_HIP(hipMemcpyAsync(&h_sum, d_sum, 8, hipMemcpyDeviceToHost, ctx->stream));
    SH_HIP(hipFreeAsync(d_sum, ctx->stream));
    SH_HIP(hipStreamSynchronize(ctx->stream));
    if (rc != STORM_HIP_OK) return rc;
    float ms = 0;
    hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    s.ms_device = ms;
    if (c->dense) {
        hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]);
        s.ms_bb = ms;
    }
    s.sum = h_sum;
    c->tri_valid = (tri != nullptr);
    c->tri_density = tri && !c->dense && o.out_mode == STORM_HIP_OUT_UPPER_TRI_DENSITY;
    s.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return STORM_HIP_OK;
}

int storm_hip_allpairs_xy(storm_hip_ctx* ctx, const storm_hip_corpus* x,
                          const storm_hip_corpus* y, uint32_t bb_kernel, uint64_t* sum,
                          uint32_t* out) {
    if (!ctx || !x || !y || x->ctx != ctx || y->ctx != ctx || !sum) return STORM_HIP_ERR_INVALID;
    *sum = 0;
    if (!x->dense || !y->dense) {
        set_error("the rectangular form takes two dense corpora");
        return STORM_HIP_ERR_UNSUPPORTED;
    }
    if (x->universe != y->universe || x->ld != y->ld) {
        set_error("rectangular corpora differ in universe (%u vs %u bits)", x->universe,
                  y->universe);
        return STORM_HIP_ERR_INVALID;
    }
    if (bb_kernel > STORM_HIP_BB_MFMA_FP4 || bb_kernel == STORM_HIP_BB_VALU) {
        set_error("the rectangular form runs on the matrix cores (AUTO, MFMA or MFMA_FP4)");
        return STORM_HIP_ERR_UNSUPPORTED;
    }
    const bool fp4 = bb_kernel == STORM_HIP_BB_MFMA_FP4 ||
                     (bb_kernel == STORM_HIP_BB_AUTO && x->n_words < (1u << 18));
    SH_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const uint64_t count = (uint64_t)x->n_rows * y->n_rows;
    uint32_t* d_out = nullptr;
    unsigned long long* d_sum = nullptr;
    if (out && count) SH_HIP(hipMalloc(&d_out, count * 4));
    hipError_t e = hipMalloc(&d_sum, 8);
    if (e == hipSuccess) e = hipMemsetAsync(d_sum, 0, 8, s);
    int rc = e == hipSuccess ? STORM_HIP_OK : STORM_HIP_ERR_DEVICE;
    if (rc == STORM_HIP_OK)
        rc = launch_bb_rect(fp4, ctx->mfma, s, x->d_words, x->n_rows, y->d_words, y->n_rows, x->ld,
                            x->n_words, d_out, d_sum, (uint32_t)ctx->cu_count);
    unsigned long long h_sum = 0;
    if (rc == STORM_HIP_OK) {
        e = hipMemcpyAsync(&h_sum, d_sum, 8, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess && d_out)
            e = hipMemcpyAsync(out, d_out, count * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            set_error("rectangular call: %s", hipGetErrorString(e));
            rc = STORM_HIP_ERR_DEVICE;
        }
    }
    (void)hipStreamSynchronize(s);
    if (d_out) (void)hipFree(d_out);
    if (d_sum) (void)hipFree(d_sum);
    if (rc == STORM_HIP_OK) *sum = h_sum;
    return rc;
}

int storm_hip_result_copy_tri(storm_hip_ctx* ctx, const storm_hip_corpus* c, uint64_t offset,
                              uint64_t count, uint32_t* out) {
    if (!ctx || !c || !c->tri_valid) return STORM_HIP_ERR_INVALID;
    if (!out && count) return STORM_HIP_ERR_NULL_VALUES;
    // Written so that no sum can wrap: offset and count may each be near 2^64.
    if (offset > c->tri_count || count > c->tri_count - offset) return STORM_HIP_ERR_RANGE;
    if (!count) return STORM_HIP_OK;
    SH_HIP(hipSetDevice(ctx->device));
    SH_HIP(hipMemcpyAsync(out, c->d_tri + offset, count * 4, hipMemcpyDeviceToHost, ctx->stream));
    SH_HIP(hipStreamSynchronize(ctx->stream));
    return STORM_HIP_OK;
}

int storm_hip_result_row_sums(storm_hip_ctx* ctx, const storm_hip_corpus* c, uint64_t* row_sum) {
    if (!ctx || !c || !c->tri_valid) return STORM_HIP_ERR_INVALID;
    if (!row_sum) return STORM_HIP_ERR_NULL_VALUES;
    if (!c->n_rows) return STORM_HIP_OK;
    SH_HIP(hipSetDevice(ctx->device));
    unsigned long long* d = nullptr;
    SH_HIP(hipMalloc(&d, (size_t)c->n_rows * 8));
    hipLaunchKernelGGL(tri_row_sums_kernel, dim3(c->n_rows), dim3(256), 0, ctx->stream, c->d_tri,
                       c->n_rows, c->tri_density ? c->d_perm : nullptr, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(row_sum, d, (size_t)c->n_rows * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(d);
    if (e != hipSuccess) {
        set_error("row sums: %s", hipGetErrorString(e));
        return STORM_HIP_ERR_DEVICE;
    }
    return STORM_HIP_OK;
}

}  // extern "C"
