// Learned per-tile cell selector: a small MLP trained on the device by this
// file's own forward, backward and Adam kernels, on timings the engine makes
// by running every eligible cell on sampled tiles.
//
// Network: F features -> H1 ReLU -> H2 ReLU -> O outputs, fp32. Output q is
// the predicted log2 of the tile's time under cell q; selection takes the
// cheapest ELIGIBLE cell. A tile whose features leave the training range
// (per feature, widened by a margin) falls back to the analytic model.
// Parameters are packed W1[H1][F] b1[H1] W2[H2][H1] b2[H2] W3[O][H2] b3[O].
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "device_common.h"
#include "general.h"

namespace stormhip {

constexpr int kF = STORM_HIP_MLP_FEATURES, kH1 = 32, kH2 = 32, kO = 10;
constexpr int kParams = kH1 * kF + kH1 + kH2 * kH1 + kH2 + kO * kH2 + kO;
constexpr int oW1 = 0, oB1 = oW1 + kH1 * kF, oW2 = oB1 + kH1, oB2 = oW2 + kH2 * kH1,
              oW3 = oB2 + kH2, oB3 = oW3 + kO * kH2;

struct Selector {
    float* params = nullptr;       // device, kParams
    float* m = nullptr;            // Adam first moment
    float* v = nullptr;            // Adam second moment
    float* grad = nullptr;
    float* norm = nullptr;         // device: mean[F], inv_std[F], lo[F], hi[F]
    std::vector<float> h_norm;     // host copy (4F)
    int step = 0;
    bool trained = false;
};

static Selector* selector_of(storm_hip_ctx* ctx, bool create) {
    Selector* s = static_cast<Selector*>(ctx->selector);
    if (s || !create) return s;
    s = new (std::nothrow) Selector();
    if (!s) return nullptr;
    if (hipMalloc(&s->params, kParams * 4) != hipSuccess || hipMalloc(&s->m, kParams * 4) != hipSuccess ||
        hipMalloc(&s->v, kParams * 4) != hipSuccess || hipMalloc(&s->grad, kParams * 4) != hipSuccess ||
        hipMalloc(&s->norm, 4 * kF * 4) != hipSuccess) {
        delete s;
        return nullptr;
    }
    s->h_norm.assign(4 * kF, 0.f);
    for (int f = 0; f < kF; ++f) { s->h_norm[kF + f] = 1.f; s->h_norm[2 * kF + f] = -1e30f; s->h_norm[3 * kF + f] = 1e30f; }
    (void)hipMemcpy(s->norm, s->h_norm.data(), 4 * kF * 4, hipMemcpyHostToDevice);
    ctx->selector = s;
    return s;
}

void selector_destroy(storm_hip_ctx* ctx) {
    Selector* s = static_cast<Selector*>(ctx->selector);
    if (!s) return;
    for (void* p : {(void*)s->params, (void*)s->m, (void*)s->v, (void*)s->grad, (void*)s->norm})
        if (p) (void)hipFree(p);
    delete s;
    ctx->selector = nullptr;
}

// ---- features ---------------------------------------------------------------
// Metadata only (no row data): the two tiles' mean/max cardinality, runs,
// non-zero words, EWAH length, bitmap residency, the universe, the span overlap
// and the tile's pair count.
__device__ inline void tile_features(const TileMeta& a, const TileMeta& b, uint32_t universe,
                                     float* x) {
    auto l2 = [](double v) { return (float)log2(1.0 + v); };
    x[0] = l2(a.card);  x[1] = l2(b.card);
    x[2] = l2(a.runs);  x[3] = l2(b.runs);
    x[4] = l2(a.nnz);   x[5] = l2(b.nnz);
    x[6] = l2(a.max_card); x[7] = l2(b.max_card);
    x[8] = l2(a.max_ewah); x[9] = l2(b.max_ewah);
    x[10] = l2(max(a.nw, b.nw));
    x[11] = (float)a.all_bitmap; x[12] = (float)b.all_bitmap;
    const double lo = max(a.first, b.first), hi = min(a.last, b.last);
    x[13] = (float)(hi >= lo ? (hi - lo + 1.0) / (double)max(1u, universe) : 0.0);
    x[14] = l2((double)a.rows * b.rows);
    x[15] = l2(a.min_card);
}

__global__ void features_kernel(const TileMeta* __restrict__ tm, uint32_t T, uint64_t P,
                                uint32_t universe, float* __restrict__ X) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P) return;
    uint32_t bi, bj;
    tri_tile((uint32_t)t, T, bi, bj);
    tile_features(tm[bi], tm[bj], universe, X + t * kF);
}

// ---- forward / backward / Adam ---------------------------------------------
__device__ inline void mlp_forward(const float* __restrict__ P, const float* x, float* a1, float* a2,
                                   float* y) {
    for (int i = 0; i < kH1; ++i) {
        float s = P[oB1 + i];
        for (int k = 0; k < kF; ++k) s = fmaf(P[oW1 + i * kF + k], x[k], s);
        a1[i] = s > 0.f ? s : 0.f;
    }
    for (int i = 0; i < kH2; ++i) {
        float s = P[oB2 + i];
        for (int k = 0; k < kH1; ++k) s = fmaf(P[oW2 + i * kH1 + k], a1[k], s);
        a2[i] = s > 0.f ? s : 0.f;
    }
    for (int i = 0; i < kO; ++i) {
        float s = P[oB3 + i];
        for (int k = 0; k < kH2; ++k) s = fmaf(P[oW3 + i * kH2 + k], a2[k], s);
        y[i] = s;
    }
}

// One thread per sample; gradients reduced in LDS, then once per block into
// `grad` (which the caller zeroes). Loss = sum over masked outputs of
// (y - t)^2 / 2, divided by `count` (the number of masked outputs in the batch).
// Forward and backward over kSB samples per workgroup: lanes 0..kSB-1 run
// one sample each and leave its activations and deltas in LDS; then every
// thread sums its parameters' gradients over the block's samples (one global
// add per parameter and block, no contended atomics).
constexpr int kSB = 64;

__global__ __launch_bounds__(256) void mlp_fwd_bwd_kernel(
    const float* __restrict__ P, const float* __restrict__ X, const float* __restrict__ Tg,
    const uint8_t* __restrict__ M, uint32_t n, float inv_count, float* __restrict__ grad,
    float* __restrict__ loss) {
    __shared__ float sx[kSB][kF], sa1[kSB][kH1], sa2[kSB][kH2];
    __shared__ float sd1[kSB][kH1], sd2[kSB][kH2], sdy[kSB][kO];
    __shared__ float lpart[kSB];
    const uint32_t s0 = blockIdx.x * kSB;
    const int t = threadIdx.x;
    if (t < kSB) {
        const uint32_t s = s0 + t;
        float x[kF], a1[kH1], a2[kH2], y[kO], dy[kO], d2[kH2];
        float l = 0.f;
        if (s < n) {
            for (int k = 0; k < kF; ++k) x[k] = X[(uint64_t)s * kF + k];
            mlp_forward(P, x, a1, a2, y);
            for (int i = 0; i < kO; ++i) {
                const bool on = M[(uint64_t)s * kO + i] != 0;
                const float e = on ? (y[i] - Tg[(uint64_t)s * kO + i]) : 0.f;
                l += 0.5f * e * e * inv_count;
                dy[i] = e * inv_count;
            }
            for (int k = 0; k < kH2; ++k) {
                float acc = 0.f;
                for (int i = 0; i < kO; ++i) acc = fmaf(P[oW3 + i * kH2 + k], dy[i], acc);
                d2[k] = a2[k] > 0.f ? acc : 0.f;
            }
            for (int k = 0; k < kH1; ++k) {
                float acc = 0.f;
                for (int i = 0; i < kH2; ++i) acc = fmaf(P[oW2 + i * kH1 + k], d2[i], acc);
                sd1[t][k] = a1[k] > 0.f ? acc : 0.f;
            }
        } else {
            for (int k = 0; k < kF; ++k) x[k] = 0.f;
            for (int k = 0; k < kH1; ++k) { a1[k] = 0.f; sd1[t][k] = 0.f; }
            for (int k = 0; k < kH2; ++k) { a2[k] = 0.f; d2[k] = 0.f; }
            for (int i = 0; i < kO; ++i) dy[i] = 0.f;
        }
        for (int k = 0; k < kF; ++k) sx[t][k] = x[k];
        for (int k = 0; k < kH1; ++k) sa1[t][k] = a1[k];
        for (int k = 0; k < kH2; ++k) { sa2[t][k] = a2[k]; sd2[t][k] = d2[k]; }
        for (int i = 0; i < kO; ++i) sdy[t][i] = dy[i];
        lpart[t] = l;
    }
    __syncthreads();
    if (t == 0 && loss) {
        float sum = 0.f;
        for (int q = 0; q < kSB; ++q) sum += lpart[q];
        atomicAdd(loss, sum);
    }
    for (int k = t; k < kParams; k += blockDim.x) {
        float acc = 0.f;
        if (k < oB1) {
            const int i = k / kF, j = k % kF;
            for (int q = 0; q < kSB; ++q) acc = fmaf(sd1[q][i], sx[q][j], acc);
        } else if (k < oW2) {
            const int i = k - oB1;
            for (int q = 0; q < kSB; ++q) acc += sd1[q][i];
        } else if (k < oB2) {
            const int i = (k - oW2) / kH1, j = (k - oW2) % kH1;
            for (int q = 0; q < kSB; ++q) acc = fmaf(sd2[q][i], sa1[q][j], acc);
        } else if (k < oW3) {
            const int i = k - oB2;
            for (int q = 0; q < kSB; ++q) acc += sd2[q][i];
        } else if (k < oB3) {
            const int i = (k - oW3) / kH2, j = (k - oW3) % kH2;
            for (int q = 0; q < kSB; ++q) acc = fmaf(sdy[q][i], sa2[q][j], acc);
        } else {
            const int i = k - oB3;
            for (int q = 0; q < kSB; ++q) acc += sdy[q][i];
        }
        if (acc != 0.f) atomicAdd(&grad[k], acc);
    }
}

// Adam (Kingma & Ba 2015, Algorithm 1), one thread per parameter.
__global__ void adam_kernel(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ m,
                            float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = G[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi; v[i] = vi;
    P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void normalize_kernel(const float* __restrict__ Xin, const float* __restrict__ norm,
                                 uint64_t n, float* __restrict__ Xout) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * kF) return;
    const int f = (int)(i % kF);
    Xout[i] = (Xin[i] - norm[f]) * norm[kF + f];
}

// Inference for every tile pair: predicted log2 time per cell, and whether the
// tile's features lie inside the training range (else analytic fallback).
__global__ __launch_bounds__(256) void mlp_infer_kernel(
    const float* __restrict__ P, const float* __restrict__ norm, const TileMeta* __restrict__ tm,
    uint32_t T, uint64_t n_pairs, uint32_t universe, float* __restrict__ cost,
    uint8_t* __restrict__ ok) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_pairs) return;
    uint32_t bi, bj;
    tri_tile((uint32_t)t, T, bi, bj);
    float x[kF], a1[kH1], a2[kH2], y[kO];
    tile_features(tm[bi], tm[bj], universe, x);
    bool in = true;
    for (int f = 0; f < kF; ++f) {
        in = in && x[f] >= norm[2 * kF + f] && x[f] <= norm[3 * kF + f];
        x[f] = (x[f] - norm[f]) * norm[kF + f];
    }
    mlp_forward(P, x, a1, a2, y);
    for (int q = 0; q < kO; ++q) cost[t * kO + q] = y[q];
    ok[t] = in ? 1 : 0;
}

__global__ void raw_forward_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                   uint32_t n, float* __restrict__ Y) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float x[kF], a1[kH1], a2[kH2], y[kO];
    for (int k = 0; k < kF; ++k) x[k] = X[(uint64_t)s * kF + k];
    mlp_forward(P, x, a1, a2, y);
    for (int q = 0; q < kO; ++q) Y[(uint64_t)s * kO + q] = y[q];
}

int learned_scores(storm_hip_ctx* ctx, const TileMeta* tm, uint32_t T, uint64_t P,
                   uint32_t universe, float* cost, uint8_t* ok, bool* have_model) {
    Selector* s = selector_of(ctx, false);
    *have_model = s && s->trained;
    if (!*have_model || P == 0) return STORM_HIP_OK;
    hipLaunchKernelGGL(mlp_infer_kernel, dim3((uint32_t)((P + 255) / 256)), dim3(256), 0,
                       ctx->stream, s->params, s->norm, tm, T, P, universe, cost, ok);
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

int launch_features(hipStream_t st, const TileMeta* tm, uint32_t T, uint64_t P, uint32_t universe,
                    float* X) {
    if (P == 0) return STORM_HIP_OK;
    hipLaunchKernelGGL(features_kernel, dim3((uint32_t)((P + 255) / 256)), dim3(256), 0, st, tm, T,
                       P, universe, X);
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

namespace {

template <class T>
struct DevBuf {
    T* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    int alloc(size_t n) {
        SH_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
        return STORM_HIP_OK;
    }
};

// One fwd/bwd over [x, t, mask] with the current params into s->grad.
int fwd_bwd(hipStream_t st, Selector* s, const float* x, const float* t, const uint8_t* mask,
            uint32_t n, uint64_t count, float* d_loss) {
    SH_HIP(hipMemsetAsync(s->grad, 0, kParams * 4, st));
    if (d_loss) SH_HIP(hipMemsetAsync(d_loss, 0, 4, st));
    if (n == 0) return STORM_HIP_OK;
    hipLaunchKernelGGL(mlp_fwd_bwd_kernel, dim3((n + kSB - 1) / kSB), dim3(256), 0, st, s->params,
                       x, t, mask, n, count ? 1.f / (float)count : 0.f, s->grad, d_loss);
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

int adam(hipStream_t st, Selector* s, float lr) {
    s->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)s->step), c2 = 1.f - std::pow(b2, (float)s->step);
    hipLaunchKernelGGL(adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, st, s->params,
                       s->grad, s->m, s->v, kParams, lr, b1, b2, 1e-8f, c1, c2);
    SH_HIP(hipGetLastError());
    return STORM_HIP_OK;
}

}  // namespace

}  // namespace stormhip

using namespace stormhip;

extern "C" {

int storm_hip_mlp_shape(int* f, int* h1, int* h2, int* o, int* n_params) {
    if (f) *f = kF;
    if (h1) *h1 = kH1;
    if (h2) *h2 = kH2;
    if (o) *o = kO;
    if (n_params) *n_params = kParams;
    return STORM_HIP_OK;
}

int storm_hip_selector_set_params(storm_hip_ctx* ctx, const float* params, const float* norm) {
    if (!ctx || !params) return STORM_HIP_ERR_INVALID;
    SH_HIP(hipSetDevice(ctx->device));
    Selector* s = selector_of(ctx, true);
    if (!s) return STORM_HIP_ERR_NOMEM;
    SH_HIP(hipMemcpy(s->params, params, kParams * 4, hipMemcpyHostToDevice));
    SH_HIP(hipMemset(s->m, 0, kParams * 4));
    SH_HIP(hipMemset(s->v, 0, kParams * 4));
    s->step = 0;
    if (norm) {
        std::memcpy(s->h_norm.data(), norm, 4 * kF * 4);
        SH_HIP(hipMemcpy(s->norm, norm, 4 * kF * 4, hipMemcpyHostToDevice));
        s->trained = true;
    }
    ++ctx->model_version;
    return STORM_HIP_OK;
}

int storm_hip_selector_get_params(storm_hip_ctx* ctx, float* params, float* norm) {
    if (!ctx) return STORM_HIP_ERR_INVALID;
    Selector* s = selector_of(ctx, false);
    if (!s) return STORM_HIP_ERR_INVALID;
    SH_HIP(hipSetDevice(ctx->device));
    if (params) SH_HIP(hipMemcpy(params, s->params, kParams * 4, hipMemcpyDeviceToHost));
    if (norm) std::memcpy(norm, s->h_norm.data(), 4 * kF * 4);
    return STORM_HIP_OK;
}

int storm_hip_selector_grad(storm_hip_ctx* ctx, const float* x, const float* t, const uint8_t* mask,
                            uint32_t n, float* grad_out, float* loss_out) {
    if (!ctx || (!x && n) || !grad_out) return STORM_HIP_ERR_INVALID;
    Selector* s = selector_of(ctx, false);
    if (!s) return STORM_HIP_ERR_INVALID;
    SH_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DevBuf<float> dx, dt, dl;
    DevBuf<uint8_t> dm;
    SH_TRY(dx.alloc((size_t)n * kF)); SH_TRY(dt.alloc((size_t)n * kO));
    SH_TRY(dm.alloc((size_t)n * kO)); SH_TRY(dl.alloc(1));
    uint64_t count = 0;
    for (uint64_t i = 0; i < (uint64_t)n * kO; ++i) count += mask[i] != 0;
    SH_HIP(hipMemcpyAsync(dx.p, x, (size_t)n * kF * 4, hipMemcpyHostToDevice, st));
    SH_HIP(hipMemcpyAsync(dt.p, t, (size_t)n * kO * 4, hipMemcpyHostToDevice, st));
    SH_HIP(hipMemcpyAsync(dm.p, mask, (size_t)n * kO, hipMemcpyHostToDevice, st));
    SH_TRY(fwd_bwd(st, s, dx.p, dt.p, dm.p, n, count, dl.p));
    SH_HIP(hipMemcpyAsync(grad_out, s->grad, kParams * 4, hipMemcpyDeviceToHost, st));
    float l = 0;
    SH_HIP(hipMemcpyAsync(&l, dl.p, 4, hipMemcpyDeviceToHost, st));
    SH_HIP(hipStreamSynchronize(st));
    if (loss_out) *loss_out = l;
    return STORM_HIP_OK;
}

int storm_hip_selector_adam_step(storm_hip_ctx* ctx, float lr) {
    if (!ctx) return STORM_HIP_ERR_INVALID;
    Selector* s = selector_of(ctx, false);
    if (!s) return STORM_HIP_ERR_INVALID;
    SH_HIP(hipSetDevice(ctx->device));
    SH_TRY(adam(ctx->stream, s, lr));
    SH_HIP(hipStreamSynchronize(ctx->stream));
    return STORM_HIP_OK;
}

int storm_hip_selector_train(storm_hip_ctx* ctx, const float* x, const float* t, const uint8_t* mask,
                             uint32_t n, uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                             float* final_loss) {
    if (!ctx || !x || !t || !mask || n == 0 || batch == 0) return STORM_HIP_ERR_INVALID;
    SH_HIP(hipSetDevice(ctx->device));
    Selector* s = selector_of(ctx, true);
    if (!s) return STORM_HIP_ERR_NOMEM;
    hipStream_t st = ctx->stream;
    // Feature normalisation and the training range (widened by a margin of
    // half a standard deviation plus 0.5, in the log2 feature space).
    std::vector<double> mean(kF, 0), var(kF, 0);
    std::vector<float> lo(kF, 1e30f), hi(kF, -1e30f);
    for (uint32_t i = 0; i < n; ++i)
        for (int f = 0; f < kF; ++f) {
            const double v = x[(uint64_t)i * kF + f];
            mean[f] += v;
            lo[f] = std::min(lo[f], (float)v);
            hi[f] = std::max(hi[f], (float)v);
        }
    for (int f = 0; f < kF; ++f) mean[f] /= n;
    for (uint32_t i = 0; i < n; ++i)
        for (int f = 0; f < kF; ++f) {
            const double d = x[(uint64_t)i * kF + f] - mean[f];
            var[f] += d * d;
        }
    for (int f = 0; f < kF; ++f) {
        const double sd = std::sqrt(var[f] / n);
        s->h_norm[f] = (float)mean[f];
        s->h_norm[kF + f] = sd > 1e-6 ? (float)(1.0 / sd) : 1.f;
        const float margin = (float)(0.5 * sd + 0.5);
        s->h_norm[2 * kF + f] = lo[f] - margin;
        s->h_norm[3 * kF + f] = hi[f] + margin;
    }
    SH_HIP(hipMemcpyAsync(s->norm, s->h_norm.data(), 4 * kF * 4, hipMemcpyHostToDevice, st));
    // He initialisation from the seeded stream.
    std::vector<float> p(kParams, 0.f);
    uint64_t rs = seed ? seed : kGolden;
    auto gauss = [&rs]() {
        rs += kGolden; const double u1 = ((splitmix_mix(rs) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        rs += kGolden; const double u2 = ((splitmix_mix(rs) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int i = 0; i < kH1 * kF; ++i) p[oW1 + i] = (float)(gauss() * std::sqrt(2.0 / kF));
    for (int i = 0; i < kH2 * kH1; ++i) p[oW2 + i] = (float)(gauss() * std::sqrt(2.0 / kH1));
    for (int i = 0; i < kO * kH2; ++i) p[oW3 + i] = (float)(gauss() * std::sqrt(1.0 / kH2));
    double tmean = 0; uint64_t tc = 0;
    for (uint64_t i = 0; i < (uint64_t)n * kO; ++i) if (mask[i]) { tmean += t[i]; ++tc; }
    for (int i = 0; i < kO; ++i) p[oB3 + i] = tc ? (float)(tmean / tc) : 0.f;
    SH_HIP(hipMemcpyAsync(s->params, p.data(), kParams * 4, hipMemcpyHostToDevice, st));
    SH_HIP(hipMemsetAsync(s->m, 0, kParams * 4, st));
    SH_HIP(hipMemsetAsync(s->v, 0, kParams * 4, st));
    s->step = 0;
    // Data on the device, normalised there; minibatches are a seeded shuffle.
    DevBuf<float> draw, dx, dt, dl;
    DevBuf<uint8_t> dm;
    SH_TRY(draw.alloc((size_t)n * kF)); SH_TRY(dx.alloc((size_t)n * kF));
    SH_TRY(dt.alloc((size_t)n * kO)); SH_TRY(dm.alloc((size_t)n * kO)); SH_TRY(dl.alloc(1));
    // Shuffle on the host (a permutation applied while uploading each epoch
    // would cost n*F copies per epoch; one fixed shuffle plus rotating batch
    // offsets is enough for this tiny model).
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    for (uint32_t i = n; i > 1; --i) {
        rs += kGolden;
        const uint32_t j = (uint32_t)(splitmix_mix(rs) % i);
        std::swap(order[i - 1], order[j]);
    }
    std::vector<float> hx((size_t)n * kF), ht((size_t)n * kO);
    std::vector<uint8_t> hm((size_t)n * kO);
    for (uint32_t i = 0; i < n; ++i) {
        std::memcpy(&hx[(size_t)i * kF], &x[(size_t)order[i] * kF], kF * 4);
        std::memcpy(&ht[(size_t)i * kO], &t[(size_t)order[i] * kO], kO * 4);
        std::memcpy(&hm[(size_t)i * kO], &mask[(size_t)order[i] * kO], kO);
    }
    SH_HIP(hipMemcpyAsync(draw.p, hx.data(), hx.size() * 4, hipMemcpyHostToDevice, st));
    SH_HIP(hipMemcpyAsync(dt.p, ht.data(), ht.size() * 4, hipMemcpyHostToDevice, st));
    SH_HIP(hipMemcpyAsync(dm.p, hm.data(), hm.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(normalize_kernel, dim3((uint32_t)(((uint64_t)n * kF + 255) / 256)), dim3(256),
                       0, st, draw.p, s->norm, (uint64_t)n, dx.p);
    SH_HIP(hipGetLastError());
    std::vector<uint64_t> cnt_prefix((size_t)n + 1, 0);
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t c = 0;
        for (int q = 0; q < kO; ++q) c += hm[(size_t)i * kO + q] != 0;
        cnt_prefix[i + 1] = cnt_prefix[i] + c;
    }
    const uint32_t bsz = std::min(batch, n);
    for (uint32_t e = 0; e < epochs; ++e) {
        for (uint32_t b0 = 0; b0 < n; b0 += bsz) {
            const uint32_t nb = std::min(bsz, n - b0);
            const uint64_t count = cnt_prefix[b0 + nb] - cnt_prefix[b0];
            SH_TRY(fwd_bwd(st, s, dx.p + (size_t)b0 * kF, dt.p + (size_t)b0 * kO,
                           dm.p + (size_t)b0 * kO, nb, count, nullptr));
            SH_TRY(adam(st, s, lr));
        }
    }
    // Final full-batch loss.
    SH_TRY(fwd_bwd(st, s, dx.p, dt.p, dm.p, n, cnt_prefix[n], dl.p));
    float l = 0;
    SH_HIP(hipMemcpyAsync(&l, dl.p, 4, hipMemcpyDeviceToHost, st));
    SH_HIP(hipStreamSynchronize(st));
    if (final_loss) *final_loss = l;
    s->trained = true;
    ++ctx->model_version;
    return STORM_HIP_OK;
}

int storm_hip_selector_predict(storm_hip_ctx* ctx, const float* x, uint32_t n, float* y) {
    if (!ctx || (!x && n) || (!y && n)) return STORM_HIP_ERR_INVALID;
    Selector* s = selector_of(ctx, false);
    if (!s) return STORM_HIP_ERR_INVALID;
    if (n == 0) return STORM_HIP_OK;
    SH_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DevBuf<float> draw, dx, dy;
    SH_TRY(draw.alloc((size_t)n * kF)); SH_TRY(dx.alloc((size_t)n * kF)); SH_TRY(dy.alloc((size_t)n * kO));
    SH_HIP(hipMemcpyAsync(draw.p, x, (size_t)n * kF * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(normalize_kernel, dim3((uint32_t)(((uint64_t)n * kF + 255) / 256)), dim3(256),
                       0, st, draw.p, s->norm, (uint64_t)n, dx.p);
    hipLaunchKernelGGL(raw_forward_kernel, dim3((n + 255) / 256), dim3(256), 0, st, s->params, dx.p,
                       n, dy.p);
    SH_HIP(hipGetLastError());
    SH_HIP(hipMemcpyAsync(y, dy.p, (size_t)n * kO * 4, hipMemcpyDeviceToHost, st));
    SH_HIP(hipStreamSynchronize(st));
    return STORM_HIP_OK;
}

}  // extern "C"
