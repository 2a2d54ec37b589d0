// Device representation building.
//
// From sorted, distinct per-row positions resident in HBM this builds what
// build_row (kernels/storm_repr.cpp:102-207) builds -- maximal runs, EWAH-64
// with the marker tables the cells index, complement lists of rows above half
// density, and the RowMeta fields with the same rank/occ gates -- with one
// wavefront per row walking its list 64 positions at a time:
//   1. count:    validation, meta, and per-row sizes of every arena;
//   2. scan:     exclusive prefix sums of the sizes -> the arenas' offsets;
//   3. fill:     the arenas, each marker written with a provisional fill bit;
//   4. markers:  fill and literal lengths of every marker from its successor.
// Words of a row are formed in registers (a segmented OR across the lanes that
// hold one word's positions), so a row costs its list once per pass whatever
// its universe; the reference's build_ewah scans the row's whole bitmap.
//
// EWAH parse (build_ewah, kernels/storm_repr.cpp:78-100): a marker is a
// maximal fill of all-zero or all-one words followed by the maximal run of
// literal words after it; a row that starts with a literal gets a marker with
// an empty fill. Hence a marker starts at every maximal zero or one run and at
// a leading literal, and marker m spans decoded words [mk_word[m],
// mk_word[m+1]) and buffer words [mk_pos[m], mk_pos[m+1]), which fixes its
// literal length (buffer span - 1) and fill length (decoded span - literals).
//
// STORMBIN (tools/sets2bin.py:23-25, loader bench/bench_allpairs.cpp:44-72):
// the row records are uploaded as they lie in the file and the positions are
// gathered into the CSR list on the device.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "build.h"
#include "device_common.h"
#include "general.h"

namespace stormhip {

namespace {

constexpr int kRowsPerBlock = 4;   // one wave per row, 256-thread blocks

__device__ inline uint64_t shfl_u64(uint64_t v, int src) {
    const uint32_t lo = __shfl((uint32_t)v, src, 64);
    const uint32_t hi = __shfl((uint32_t)(v >> 32), src, 64);
    return ((uint64_t)hi << 32) | lo;
}

__device__ inline uint64_t shfl_down_u64(uint64_t v, unsigned d) {
    const uint32_t lo = __shfl_down((uint32_t)v, d, 64);
    const uint32_t hi = __shfl_down((uint32_t)(v >> 32), d, 64);
    return ((uint64_t)hi << 32) | lo;
}

__device__ inline int top_lane(uint64_t m) { return 63 - __clzll((long long)m); }

// Output arrays at the row's base (fill pass only).
struct RowOut {
    uint32_t* rs; uint32_t* re;          // runs
    uint64_t* ew;                        // EWAH buffer
    uint32_t* mkw; uint32_t* mkp;        // marker tables
};

struct RowTotals {
    uint32_t runs = 0, nnz = 0, items = 0, markers = 0;
    bool bad = false;
};

// Walks one row (all 64 lanes of the wave, wave-uniform control flow).
// Counting pass: WRITE false, `o` unused.
template <bool WRITE>
__device__ RowTotals walk_row(const uint32_t* __restrict__ p, uint32_t c, uint32_t universe,
                              uint32_t nw, bool runs, bool ewah, const RowOut& o) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1;
    RowTotals t;
    bool have_prev = false, prev_ones = false;   // last completed nonzero word
    uint32_t pw = 0;
    bool have_cur = false;                       // word whose positions are still arriving
    uint32_t cw = 0;
    uint64_t cv = 0;
    uint32_t prev_x = 0;

    // One completed nonzero word, wave-uniform; lane 0 writes.
    auto emit_one = [&](uint32_t w, uint64_t v) {
        const bool ones = v == ~0ull;
        const bool gap = have_prev ? w > pw + 1 : w > 0;
        const bool start_ones = ones && (gap || !have_prev || !prev_ones);
        const bool lead_lit = !ones && !have_prev && w == 0;
        if (WRITE && lane == 0) {
            uint32_t q = t.items, m = t.markers;
            if (gap) { o.mkw[m] = have_prev ? pw + 1 : 0u; o.mkp[m] = q; o.ew[q] = 0; ++q; ++m; }
            if (start_ones) { o.mkw[m] = w; o.mkp[m] = q; o.ew[q] = 1; ++q; ++m; }
            if (lead_lit) { o.mkw[m] = 0; o.mkp[m] = q; o.ew[q] = 0; ++q; ++m; }
            if (!ones) o.ew[q] = v;
        }
        t.items += (uint32_t)gap + start_ones + lead_lit + !ones;
        t.markers += (uint32_t)gap + start_ones + lead_lit;
        pw = w; prev_ones = ones; have_prev = true;
    };

    for (uint32_t base = 0; base < c; base += 64) {
        const uint32_t cnt = min(64u, c - base);
        const bool valid = lane < cnt;
        const uint32_t x = valid ? p[base + lane] : 0u;
        uint32_t xp = __shfl_up(x, 1u, 64);
        if (lane == 0) xp = prev_x;
        const bool has_prev = base + lane > 0;
        if (__ballot(valid && (x >= universe || (has_prev && x <= xp)))) {
            t.bad = true;
            return t;
        }
        // Runs: [first, last + 1) of maximal stretches of consecutive positions.
        uint32_t xn = __shfl_down(x, 1u, 64);
        if (lane == 63) xn = (base + 64 < c) ? p[base + 64] : 0u;
        const bool has_next = base + lane + 1 < c;
        const bool rstart = valid && (!has_prev || x != xp + 1);
        const bool rend = valid && (!has_next || xn != x + 1);
        const uint64_t ms = __ballot(rstart);
        if (WRITE && runs) {
            if (rstart) o.rs[t.runs + __popcll(ms & lt)] = x;
            if (rend) o.re[t.runs + __popcll(ms & (lt | (1ull << lane))) - 1] = x + 1;
        }
        t.runs += __popcll(ms);
        // Nonzero words: a head lane starts each word present in this chunk.
        const bool wstart = valid && (!has_prev || (x >> 6) != (xp >> 6));
        const uint64_t heads = __ballot(wstart);
        t.nnz += __popcll(heads);
        if (!ewah) {
            prev_x = __shfl(x, (int)cnt - 1, 64);
            continue;
        }
        // OR of each lane's bit with the bits of the later lanes of its word.
        const uint64_t bounds = heads | (cnt < 64 ? (~0ull << cnt) : 0ull);
        const uint64_t above = lane == 63 ? 0ull : (bounds & (~0ull << (lane + 1)));
        const uint32_t seg_end = above ? (uint32_t)__ffsll((long long)above) - 1 : 64u;
        uint64_t val = valid ? (1ull << (x & 63)) : 0ull;
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint64_t y = shfl_down_u64(val, d);
            if (lane + d < seg_end) val |= y;
        }
        // Lanes before the first head continue the word carried in.
        if (!(heads & 1ull)) cv |= shfl_u64(val, 0);
        if (heads) {
            if (have_cur) emit_one(cw, cv);
            // Every head but the last completes its word here.
            const int last = top_lane(heads);
            const bool em = wstart && (int)lane != last;
            const uint32_t w = x >> 6;
            const bool ones = val == ~0ull;
            const uint64_t below = heads & lt;
            const int ph = below ? top_lane(below) : 0;
            const uint32_t w_before = __shfl(w, ph, 64);
            const uint64_t v_before = shfl_u64(val, ph);
            const bool hp = below ? true : have_prev;
            const uint32_t prevw = below ? w_before : pw;
            const bool prev_o = below ? (v_before == ~0ull) : prev_ones;
            const bool gap = em && (hp ? w > prevw + 1 : w > 0);
            const bool start_ones = em && ones && (gap || !hp || !prev_o);
            const bool lead_lit = em && !ones && !hp && w == 0;
            const bool lit = em && !ones;
            const uint64_t bg = __ballot(gap), bo = __ballot(start_ones);
            const uint64_t bl0 = __ballot(lead_lit), bl = __ballot(lit);
            if (WRITE) {
                const uint32_t mk_before =
                    __popcll(bg & lt) + __popcll(bo & lt) + __popcll(bl0 & lt);
                uint32_t q = t.items + mk_before + __popcll(bl & lt);
                uint32_t m = t.markers + mk_before;
                if (gap) { o.mkw[m] = hp ? prevw + 1 : 0u; o.mkp[m] = q; o.ew[q] = 0; ++q; ++m; }
                if (start_ones) { o.mkw[m] = w; o.mkp[m] = q; o.ew[q] = 1; ++q; ++m; }
                if (lead_lit) { o.mkw[m] = 0; o.mkp[m] = q; o.ew[q] = 0; ++q; ++m; }
                if (lit) o.ew[q] = val;
            }
            const uint32_t nmk = __popcll(bg) + __popcll(bo) + __popcll(bl0);
            t.items += nmk + __popcll(bl);
            t.markers += nmk;
            const uint64_t done = heads & ((1ull << last) - 1);
            if (done) {
                const int lh = top_lane(done);
                pw = __shfl(w, lh, 64);
                prev_ones = shfl_u64(val, lh) == ~0ull;
                have_prev = true;
            }
            cw = __shfl(w, last, 64);
            cv = shfl_u64(val, last);
            have_cur = true;
        }
        prev_x = __shfl(x, (int)cnt - 1, 64);
    }
    if (ewah) {
        if (have_cur) emit_one(cw, cv);
        // Trailing zero fill (the whole row when it has no set bit).
        if (have_prev ? nw > pw + 1 : nw > 0) {
            if (WRITE && lane == 0) {
                o.mkw[t.markers] = have_prev ? pw + 1 : 0u;
                o.mkp[t.markers] = t.items;
                o.ew[t.items] = 0;
            }
            ++t.items;
            ++t.markers;
        }
    }
    return t;
}

// Complement of a row with 2 * c > universe: position q (not in the row) sits
// at index q - (number of row positions below q). Gaps of up to 16 positions
// are written by their lane, longer ones by the whole wave.
__device__ void write_complement(const uint32_t* __restrict__ p, uint32_t c, uint32_t universe,
                                 uint32_t* __restrict__ comp) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t first = p[0];
    for (uint32_t q = lane; q < first; q += 64) comp[q] = q;
    for (uint32_t base = 0; base < c; base += 64) {
        const uint32_t i = base + lane;
        const bool valid = i < c;
        const uint32_t x = valid ? p[i] : 0u;
        const uint32_t nxt = valid ? (i + 1 < c ? p[i + 1] : universe) : 0u;
        const uint32_t g = valid ? nxt - x - 1 : 0u;   // gap (x, nxt)
        if (g <= 16)
            for (uint32_t q = x + 1; q < nxt; ++q) comp[q - i - 1] = q;
        uint64_t longs = __ballot(g > 16);
        while (longs) {
            const int s = __ffsll((long long)longs) - 1;
            longs &= longs - 1;
            const uint32_t sx = __shfl(x, s, 64), sn = __shfl(nxt, s, 64);
            const uint32_t si = base + (uint32_t)s;
            for (uint32_t q = sx + 1 + lane; q < sn; q += 64) comp[q - si - 1] = q;
        }
    }
}

__global__ __launch_bounds__(256) void repr_count_kernel(
    const uint64_t* __restrict__ offs, const uint32_t* __restrict__ pos, uint32_t n,
    uint32_t universe, uint32_t flags, uint32_t* __restrict__ meta, uint32_t* __restrict__ cnt,
    unsigned int* __restrict__ bad) {
    const uint32_t r = blockIdx.x * kRowsPerBlock + (threadIdx.x >> 6);
    if (r >= n) return;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t o0 = offs[r], o1 = offs[r + 1];
    if (o1 < o0 || o1 - o0 > 0xFFFFFFFFull) {
        if (lane == 0) atomicOr(bad, 1u);
        return;
    }
    const uint32_t c = (uint32_t)(o1 - o0);
    const uint32_t nw = (universe + 63) / 64;
    const bool runs = flags & STORM_HIP_BUILD_RUNS, ewah = flags & STORM_HIP_BUILD_EWAH;
    const RowTotals t = walk_row<false>(pos + o0, c, universe, nw, runs, ewah, RowOut{});
    if (lane != 0) return;
    if (t.bad) {
        atomicOr(bad, 1u);
        return;
    }
    uint32_t* m = meta + (uint64_t)r * 8;
    m[0] = c;
    m[1] = t.runs;
    m[2] = nw;
    m[3] = t.nnz;
    m[4] = c ? pos[o0] : 0xFFFFFFFFu;
    m[5] = c ? pos[o1 - 1] : 0u;
    const uint32_t mean_rl = t.runs ? c / t.runs : 0u;
    m[6] = mean_rl >= 512;                                  // rank gate
    m[7] = (uint64_t)c * 512 > (uint64_t)universe;          // occ gate
    const bool comp = (flags & STORM_HIP_BUILD_COMP) && (uint64_t)c * 2 > universe;
    cnt[r] = runs ? t.runs : 0u;
    cnt[(uint64_t)n + r] = ewah ? t.items : 0u;
    cnt[2 * (uint64_t)n + r] = ewah ? t.markers : 0u;
    cnt[3 * (uint64_t)n + r] = comp ? universe - c : 0u;
}

struct ScanOut { uint64_t* o[4]; };

// Exclusive scan of n u32 sizes into n + 1 u64 offsets; block k scans array k.
__global__ __launch_bounds__(1024) void scan_sizes_kernel(const uint32_t* __restrict__ cnt,
                                                          uint32_t n, ScanOut outs) {
    __shared__ uint64_t part[1024];
    const uint32_t* a = cnt + (uint64_t)blockIdx.x * n;
    uint64_t* out = outs.o[blockIdx.x];
    const uint32_t tid = threadIdx.x;
    const uint64_t per = ((uint64_t)n + 1023) / 1024;
    const uint64_t lo = tid * per, hi = min((uint64_t)n, lo + per);
    uint64_t s = 0;
    for (uint64_t i = lo; i < hi; ++i) s += a[i];
    part[tid] = s;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {   // inclusive Hillis-Steele scan
        const uint64_t y = tid >= d ? part[tid - d] : 0ull;
        __syncthreads();
        part[tid] += y;
        __syncthreads();
    }
    uint64_t acc = part[tid] - s;
    for (uint64_t i = lo; i < hi; ++i) { out[i] = acc; acc += a[i]; }
    if (tid == 1023) out[n] = part[1023];
}

__global__ __launch_bounds__(256) void repr_fill_kernel(
    const uint64_t* __restrict__ offs, const uint32_t* __restrict__ pos, uint32_t n,
    uint32_t universe, uint32_t flags, const uint64_t* __restrict__ run_off,
    const uint64_t* __restrict__ ewah_off, const uint64_t* __restrict__ mk_off,
    const uint64_t* __restrict__ comp_off, uint32_t* __restrict__ run_start,
    uint32_t* __restrict__ run_end, uint64_t* __restrict__ ewah, uint32_t* __restrict__ mk_word,
    uint32_t* __restrict__ mk_pos, uint32_t* __restrict__ comp) {
    const uint32_t r = blockIdx.x * kRowsPerBlock + (threadIdx.x >> 6);
    if (r >= n) return;
    const uint64_t o0 = offs[r];
    const uint32_t c = (uint32_t)(offs[r + 1] - o0);
    const uint32_t nw = (universe + 63) / 64;
    const bool runs = flags & STORM_HIP_BUILD_RUNS, ew = flags & STORM_HIP_BUILD_EWAH;
    if (runs || ew) {
        RowOut o;
        o.rs = run_start + run_off[r];
        o.re = run_end + run_off[r];
        o.ew = ewah + ewah_off[r];
        o.mkw = mk_word + mk_off[r];
        o.mkp = mk_pos + mk_off[r];
        (void)walk_row<true>(pos + o0, c, universe, nw, runs, ew, o);
    }
    if (comp_off[r + 1] > comp_off[r]) write_complement(pos + o0, c, universe, comp + comp_off[r]);
}

// Marker m of a row: literal length = buffer span - 1, fill length = decoded
// span - literal length, fill bit as provisionally written.
__global__ __launch_bounds__(256) void ewah_markers_kernel(
    uint32_t n, uint32_t nw, const uint64_t* __restrict__ ewah_off,
    const uint64_t* __restrict__ mk_off, const uint32_t* __restrict__ mk_word,
    const uint32_t* __restrict__ mk_pos, uint64_t* __restrict__ ewah) {
    const uint32_t r = blockIdx.x * kRowsPerBlock + (threadIdx.x >> 6);
    if (r >= n) return;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t m0 = mk_off[r];
    const uint32_t M = (uint32_t)(mk_off[r + 1] - m0);
    const uint32_t items = (uint32_t)(ewah_off[r + 1] - ewah_off[r]);
    uint64_t* buf = ewah + ewah_off[r];
    for (uint32_t m = lane; m < M; m += 64) {
        const uint32_t q = mk_pos[m0 + m], w = mk_word[m0 + m];
        const uint32_t nq = m + 1 < M ? mk_pos[m0 + m + 1] : items;
        const uint32_t nword = m + 1 < M ? mk_word[m0 + m + 1] : nw;
        const uint64_t ll = nq - q - 1;
        const uint64_t fl = nword - w - ll;
        buf[q] = (buf[q] & 1ull) | (fl << 1) | (ll << 32);
    }
}

// Positions of row r: raw[src[r] - base ...) -> list[offs[r] ...).
__global__ __launch_bounds__(256) void gather_rows_kernel(
    const uint32_t* __restrict__ raw, uint64_t raw_base, const uint64_t* __restrict__ src,
    const uint64_t* __restrict__ offs, uint32_t r0, uint32_t r1, uint32_t* __restrict__ list) {
    const uint32_t r = r0 + blockIdx.x * kRowsPerBlock + (threadIdx.x >> 6);
    if (r >= r1) return;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t c = offs[r + 1] - offs[r];
    const uint32_t* s = raw + (src[r] - raw_base);
    uint32_t* d = list + offs[r];
    for (uint64_t i = lane; i < c; i += 64) d[i] = s[i];
}

uint32_t row_blocks(uint64_t rows) {
    return (uint32_t)((rows + kRowsPerBlock - 1) / kRowsPerBlock);
}

template <class T>
int dev_alloc(T*& d, uint64_t count, uint64_t& bytes) {
    const size_t b = (count ? count : 1) * sizeof(T);
    SH_HIP(hipMalloc((void**)&d, b));
    bytes += b;
    return STORM_HIP_OK;
}

struct DevTemp {
    void* p = nullptr;
    ~DevTemp() { if (p) (void)hipFree(p); }
};

// Memory-mapped STORMBIN file and the records it keeps (offsets in u32 words
// from the start of the file).
struct StormbinScan {
    const uint32_t* words = nullptr;
    size_t map_bytes = 0;
    uint32_t universe = 0;
    std::vector<uint64_t> offs, src;   // CSR offsets, record data offsets
    ~StormbinScan() { if (words) munmap((void*)words, map_bytes); }
};

int stormbin_scan(const char* path, uint32_t max_rows, uint32_t stride, StormbinScan& S) {
    if (!path) return STORM_HIP_ERR_INVALID;
    const int fd = open(path, O_RDONLY);
    if (fd < 0) { set_error("cannot open %s", path); return STORM_HIP_ERR_INVALID; }
    struct stat st;
    if (fstat(fd, &st) != 0 || st.st_size < 24) {
        close(fd);
        set_error("%s: not a STORMBIN file", path);
        return STORM_HIP_ERR_INVALID;
    }
    S.map_bytes = (size_t)st.st_size;
    void* m = mmap(nullptr, S.map_bytes, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (m == MAP_FAILED) { set_error("cannot map %s", path); return STORM_HIP_ERR_INVALID; }
    S.words = (const uint32_t*)m;
    if (std::memcmp(m, "STORMBIN", 8) != 0) {
        set_error("%s: bad magic", path);
        return STORM_HIP_ERR_INVALID;
    }
    // Header: magic, version, n_rows, n_bits, pad (the loader uses n_bits only).
    S.universe = S.words[4];
    if (S.universe == 0) { set_error("%s: zero universe", path); return STORM_HIP_ERR_INVALID; }
    const uint64_t total = S.map_bytes / 4;
    if (stride == 0) stride = 1;
    S.offs.assign(1, 0);
    S.src.clear();
    uint64_t at = 6, seen = 0;
    while (max_rows == 0 || S.src.size() < max_rows) {
        if (at + 1 > total) break;
        const uint32_t cnt = S.words[at];
        if (at + 1 + (uint64_t)cnt > total) break;   // truncated record ends the file
        if (seen % stride == 0 && cnt > 0) {
            S.src.push_back(at + 1);
            S.offs.push_back(S.offs.back() + cnt);
        }
        ++seen;
        at += 1 + (uint64_t)cnt;
    }
    if (S.src.size() < 2) {
        set_error("%s: fewer than 2 non-empty rows", path);
        return STORM_HIP_ERR_INVALID;
    }
    if (S.src.size() > 0xFFFFFFFFull) return STORM_HIP_ERR_RANGE;
    return STORM_HIP_OK;
}

}  // namespace

int device_build(storm_hip_ctx* ctx, storm_hip_corpus* c, uint32_t flags) {
    const uint32_t n = c->n_rows;
    hipStream_t s = ctx->stream;
    uint64_t& B = c->device_bytes;
    SH_TRY(dev_alloc(c->d_meta, (uint64_t)n * 8, B));
    SH_TRY(dev_alloc(c->d_run_off, (uint64_t)n + 1, B));
    SH_TRY(dev_alloc(c->d_ewah_off, (uint64_t)n + 1, B));
    SH_TRY(dev_alloc(c->d_mk_off, (uint64_t)n + 1, B));
    SH_TRY(dev_alloc(c->d_comp_off, (uint64_t)n + 1, B));
    DevTemp cnt, bad;
    SH_HIP(hipMalloc(&cnt.p, std::max<uint64_t>(1, (uint64_t)n * 4) * 4));
    SH_HIP(hipMalloc(&bad.p, 4));
    SH_HIP(hipMemsetAsync(bad.p, 0, 4, s));
    if (n) {
        hipLaunchKernelGGL(repr_count_kernel, dim3(row_blocks(n)), dim3(256), 0, s, c->d_list_off,
                           c->d_list, n, c->universe, flags, c->d_meta, (uint32_t*)cnt.p,
                           (unsigned int*)bad.p);
        SH_HIP(hipGetLastError());
    }
    ScanOut so{{c->d_run_off, c->d_ewah_off, c->d_mk_off, c->d_comp_off}};
    hipLaunchKernelGGL(scan_sizes_kernel, dim3(4), dim3(1024), 0, s, (const uint32_t*)cnt.p, n,
                       so);
    SH_HIP(hipGetLastError());
    uint64_t tot[4];
    uint32_t bad_h = 0;
    for (int k = 0; k < 4; ++k)
        SH_HIP(hipMemcpyAsync(&tot[k], so.o[k] + n, 8, hipMemcpyDeviceToHost, s));
    SH_HIP(hipMemcpyAsync(&bad_h, bad.p, 4, hipMemcpyDeviceToHost, s));
    c->h_meta.resize((size_t)n * 8);
    if (n)
        SH_HIP(hipMemcpyAsync(c->h_meta.data(), c->d_meta, (size_t)n * 32, hipMemcpyDeviceToHost, s));
    SH_HIP(hipStreamSynchronize(s));
    if (bad_h) {
        set_error("row positions must be sorted, distinct and < universe");
        return STORM_HIP_ERR_INVALID;
    }
    SH_TRY(dev_alloc(c->d_run_start, tot[0], B));
    SH_TRY(dev_alloc(c->d_run_end, tot[0], B));
    SH_TRY(dev_alloc(c->d_ewah, tot[1], B));
    SH_TRY(dev_alloc(c->d_mk_word, tot[2], B));
    SH_TRY(dev_alloc(c->d_mk_pos, tot[2], B));
    SH_TRY(dev_alloc(c->d_comp, tot[3], B));
    if (n) {
        hipLaunchKernelGGL(repr_fill_kernel, dim3(row_blocks(n)), dim3(256), 0, s, c->d_list_off,
                           c->d_list, n, c->universe, flags, c->d_run_off, c->d_ewah_off,
                           c->d_mk_off, c->d_comp_off, c->d_run_start, c->d_run_end, c->d_ewah,
                           c->d_mk_word, c->d_mk_pos, c->d_comp);
        SH_HIP(hipGetLastError());
        if (flags & STORM_HIP_BUILD_EWAH) {
            hipLaunchKernelGGL(ewah_markers_kernel, dim3(row_blocks(n)), dim3(256), 0, s, n,
                               c->n_words, c->d_ewah_off, c->d_mk_off, c->d_mk_word, c->d_mk_pos,
                               c->d_ewah);
            SH_HIP(hipGetLastError());
        }
    }
    // The temporaries are freed when this returns: finish their users first.
    SH_HIP(hipStreamSynchronize(s));
    return STORM_HIP_OK;
}

}  // namespace stormhip

using namespace stormhip;

extern "C" {

int storm_hip_corpus_from_device_lists(storm_hip_ctx* ctx, const uint64_t* d_offs,
                                       const uint32_t* d_pos, uint32_t n_rows,
                                       uint32_t universe, uint32_t flags,
                                       storm_hip_corpus** out) {
    if (!ctx || !out) return STORM_HIP_ERR_INVALID;
    *out = nullptr;
    if (!d_offs) return STORM_HIP_ERR_NULL_VALUES;
    if (universe == 0 || (flags & 3u) == 3u) return STORM_HIP_ERR_INVALID;
    SH_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    uint64_t first = 0, last = 0;
    SH_HIP(hipMemcpy(&first, d_offs, 8, hipMemcpyDeviceToHost));
    SH_HIP(hipMemcpy(&last, d_offs + n_rows, 8, hipMemcpyDeviceToHost));
    if (last < first) return STORM_HIP_ERR_INVALID;
    if (!d_pos && last > first) return STORM_HIP_ERR_NULL_VALUES;
    storm_hip_corpus* c = corpus_new(ctx, n_rows, universe, flags);
    if (!c) return STORM_HIP_ERR_NOMEM;
    auto fail = [&](int st) { storm_hip_corpus_free(c); return st; };
    // The corpus owns copies of the lists, rebased to offset 0.
    std::vector<uint64_t> h_offs((size_t)n_rows + 1);
    if (hipMemcpy(h_offs.data(), d_offs, h_offs.size() * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(STORM_HIP_ERR_DEVICE);
    for (auto& x : h_offs) x -= first;
    for (uint32_t r = 0; r < n_rows; ++r)
        if (h_offs[r + 1] < h_offs[r]) return fail(STORM_HIP_ERR_INVALID);
    int st = upload_vec(c->d_list_off, h_offs, s, c->device_bytes);
    if (st != STORM_HIP_OK) return fail(st);
    const uint64_t total = last - first;
    if (hipMalloc((void**)&c->d_list, std::max<uint64_t>(total, 1) * 4) != hipSuccess)
        return fail(STORM_HIP_ERR_NOMEM);
    c->device_bytes += std::max<uint64_t>(total, 1) * 4;
    if (total &&
        hipMemcpyAsync(c->d_list, d_pos + first, total * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(STORM_HIP_ERR_DEVICE);
    st = device_build(ctx, c, flags);
    if (st == STORM_HIP_OK)
        st = corpus_finish(ctx, c, bitmap_rows(c->h_meta, n_rows, universe, flags & 3u));
    if (st != STORM_HIP_OK) return fail(st);
    *out = c;
    return STORM_HIP_OK;
}

int storm_hip_stormbin_read(const char* path, uint32_t max_rows, uint32_t stride,
                            uint64_t** offs, uint32_t** pos, uint32_t* n_rows,
                            uint32_t* universe) {
    if (!offs || !pos || !n_rows || !universe) return STORM_HIP_ERR_INVALID;
    *offs = nullptr;
    *pos = nullptr;
    StormbinScan S;
    SH_TRY(stormbin_scan(path, max_rows, stride, S));
    const uint32_t n = (uint32_t)S.src.size();
    uint64_t* o = (uint64_t*)std::malloc(((size_t)n + 1) * 8);
    uint32_t* p = (uint32_t*)std::malloc(std::max<uint64_t>(S.offs[n], 1) * 4);
    if (!o || !p) {
        std::free(o);
        std::free(p);
        return STORM_HIP_ERR_NOMEM;
    }
    std::memcpy(o, S.offs.data(), ((size_t)n + 1) * 8);
    for (uint32_t r = 0; r < n; ++r)
        std::memcpy(p + S.offs[r], S.words + S.src[r], (S.offs[r + 1] - S.offs[r]) * 4);
    *offs = o;
    *pos = p;
    *n_rows = n;
    *universe = S.universe;
    return STORM_HIP_OK;
}

int storm_hip_corpus_from_stormbin(storm_hip_ctx* ctx, const char* path, uint32_t max_rows,
                                   uint32_t stride, uint32_t flags, storm_hip_corpus** out) {
    if (!ctx || !out) return STORM_HIP_ERR_INVALID;
    *out = nullptr;
    if ((flags & 3u) == 3u) return STORM_HIP_ERR_INVALID;
    StormbinScan S;
    SH_TRY(stormbin_scan(path, max_rows, stride, S));
    SH_HIP(hipSetDevice(ctx->device));
    const uint32_t n = (uint32_t)S.src.size();
    storm_hip_corpus* c = corpus_new(ctx, n, S.universe, flags);
    if (!c) return STORM_HIP_ERR_NOMEM;
    auto fail = [&](int st) { storm_hip_corpus_free(c); return st; };
    hipStream_t s = ctx->stream;
    int st = upload_vec(c->d_list_off, S.offs, s, c->device_bytes);
    if (st != STORM_HIP_OK) return fail(st);
    const uint64_t total = S.offs[n];
    if (hipMalloc((void**)&c->d_list, std::max<uint64_t>(total, 1) * 4) != hipSuccess)
        return fail(STORM_HIP_ERR_NOMEM);
    c->device_bytes += std::max<uint64_t>(total, 1) * 4;
    uint64_t* d_src = nullptr;
    st = upload_vec(d_src, S.src, s, c->device_bytes);
    DevTemp src_guard;
    src_guard.p = d_src;
    c->device_bytes -= std::max<size_t>(S.src.size(), 1) * 8;
    if (st != STORM_HIP_OK) return fail(st);
    // Records go up in chunks of about 64 Mi words (a longer row is a chunk of
    // its own); each chunk is gathered into the list before the next lands.
    constexpr uint64_t kChunkWords = 64ull << 20;
    uint64_t max_span = 0;
    for (uint32_t r0 = 0; r0 < n;) {
        uint32_t r1 = r0 + 1;
        while (r1 < n && S.src[r1] + (S.offs[r1 + 1] - S.offs[r1]) - S.src[r0] <= kChunkWords) ++r1;
        max_span = std::max(max_span, S.src[r1 - 1] + (S.offs[r1] - S.offs[r1 - 1]) - S.src[r0]);
        r0 = r1;
    }
    DevTemp raw;
    if (hipMalloc(&raw.p, max_span * 4) != hipSuccess) return fail(STORM_HIP_ERR_NOMEM);
    for (uint32_t r0 = 0; r0 < n;) {
        uint32_t r1 = r0 + 1;
        while (r1 < n && S.src[r1] + (S.offs[r1 + 1] - S.offs[r1]) - S.src[r0] <= kChunkWords) ++r1;
        const uint64_t span = S.src[r1 - 1] + (S.offs[r1] - S.offs[r1 - 1]) - S.src[r0];
        if (hipMemcpyAsync(raw.p, S.words + S.src[r0], span * 4, hipMemcpyHostToDevice, s) !=
            hipSuccess)
            return fail(STORM_HIP_ERR_DEVICE);
        hipLaunchKernelGGL(gather_rows_kernel, dim3(row_blocks(r1 - r0)), dim3(256), 0, s,
                           (const uint32_t*)raw.p, S.src[r0], d_src, c->d_list_off, r0, r1,
                           c->d_list);
        if (hipGetLastError() != hipSuccess) return fail(STORM_HIP_ERR_DEVICE);
        // The staging buffer is reused by the next chunk.
        if (hipStreamSynchronize(s) != hipSuccess) return fail(STORM_HIP_ERR_DEVICE);
        r0 = r1;
    }
    st = device_build(ctx, c, flags);
    if (st == STORM_HIP_OK)
        st = corpus_finish(ctx, c, bitmap_rows(c->h_meta, n, S.universe, flags & 3u));
    if (st != STORM_HIP_OK) return fail(st);
    *out = c;
    return STORM_HIP_OK;
}

}  // extern "C"
