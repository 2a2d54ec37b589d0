// storm_hip_allpairs_contig_noalloc: the dense all-pairs call for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops). The caller passes the rows and the outputs in device
// memory; the only other device memory the call uses is `scratch`, a
// file-scope array holding the row map of the matrix-core B x B (tile
// positions past the last row are mapped onto row 0, whose counts the
// kernel's epilogue discards, so no padding rows are needed). Each GPU keeps
// its own copy of the array until the process ends; its address on a device
// is looked up once, when the first context on that device is created.
//
// The array is shared by every call on its GPU and keeps the last call's
// contents: calls that overlap on one GPU (two contexts, two host threads)
// race on it. Calls on one context are stream-ordered and do not.
#include <cstring>
#include <mutex>

#include "device_common.h"
#include "internal.h"

namespace stormhip {

namespace {

// Row map entries for up to kNoAllocMaxRows rows plus a tile of padding.
constexpr uint32_t kNoAllocMaxRows = 1u << 20;
constexpr uint32_t kScratchEntries = kNoAllocMaxRows + kRowPad;

static __device__ __align__(256) uint32_t scratch[kScratchEntries];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

__global__ void noalloc_rowmap_kernel(uint32_t* __restrict__ rowmap, uint32_t n, uint32_t npad) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < npad) rowmap[p] = p < n ? p : 0u;
}

__global__ void noalloc_zero_sum_kernel(unsigned long long* sum) { *sum = 0; }

}  // namespace

// Looks up this device's copy of `scratch` (the device must be current).
int noalloc_register_device(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("device %d beyond the %d devices the scratch table holds", device, kMaxDevices);
        return STORM_HIP_ERR_RANGE;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return STORM_HIP_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error("hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        return STORM_HIP_ERR_DEVICE;
    }
    g_scratch_ptr[device] = static_cast<uint32_t*>(p);
    return STORM_HIP_OK;
}

}  // namespace stormhip

using namespace stormhip;

extern "C" {

int storm_hip_allpairs_contig_noalloc(storm_hip_ctx* ctx, const uint64_t* d_words,
                                      uint32_t n_rows, uint32_t n_words, uint32_t ld,
                                      uint32_t bb_kernel, uint32_t* d_tri,
                                      unsigned long long* d_sum) {
    if (!ctx || !d_sum) return STORM_HIP_ERR_INVALID;
    if (!d_words && n_rows && n_words) return STORM_HIP_ERR_NULL_VALUES;
    if (bb_kernel != STORM_HIP_BB_AUTO && bb_kernel != STORM_HIP_BB_MFMA &&
        bb_kernel != STORM_HIP_BB_MFMA_FP4) {
        set_error("the allocation-free call runs the matrix-core B x B (AUTO, MFMA, MFMA_FP4)");
        return STORM_HIP_ERR_UNSUPPORTED;
    }
    if (n_rows > kNoAllocMaxRows) {
        set_error("allocation-free call: %u rows exceed the scratch row map's %u", n_rows,
                  kNoAllocMaxRows);
        return STORM_HIP_ERR_RANGE;
    }
    if (ld < n_words || ld % 8 != 0) {
        set_error("allocation-free call: row stride %u must be >= n_words (%u) and a multiple of 8",
                  ld, n_words);
        return STORM_HIP_ERR_INVALID;
    }
    uint32_t* rowmap = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices) rowmap = g_scratch_ptr[ctx->device];
    }
    if (!rowmap) {
        set_error("allocation-free call: no scratch registered for device %d", ctx->device);
        return STORM_HIP_ERR_DEVICE;
    }
    SH_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(noalloc_zero_sum_kernel, dim3(1), dim3(1), 0, s, d_sum);
    SH_HIP(hipGetLastError());
    if (n_rows < 2) return STORM_HIP_OK;
    // Every tile position the kernel reads: ceil(n / 256) tiles of 256 rows.
    const uint32_t npad = (n_rows + kRowPad - 1) / kRowPad * kRowPad;
    hipLaunchKernelGGL(noalloc_rowmap_kernel, dim3((npad + 255) / 256), dim3(256), 0, s, rowmap,
                       n_rows, npad);
    SH_HIP(hipGetLastError());
    // No scratch allocation: K slices (small corpora) add into the zeroed
    // triangle.
    return launch_bb_dense(bb_kernel, ctx->mfma, s, d_words, ld, n_words, rowmap, n_rows, d_tri,
                           d_sum, (uint32_t)ctx->cu_count, nullptr, false);
}

}  // extern "C"
