// storm_hip_corpus_upload: general corpora from per-row arenas in the
// storm_repr layouts (kernels/storm_repr.h:112-276), taken as given.
//
// Each arena the caller supplies is checked against the row it describes and
// uploaded unchanged: sorted distinct lists, maximal runs, EWAH-64 buffers
// (whose marker tables are parsed here), complement lists, bitmaps, rank9 and
// zone maps, and RowMeta. A row needs one of a list, runs or a bitmap; lists
// are not required (rows held only as runs, or only as bitmaps, are
// corpora of their own kind: the cells that read a list are then not
// eligible). rank9 and zone maps are rebuilt on the device from the bitmaps
// (same bytes; a supplied index must equal build_rank / build_occ of its
// bitmap). The checks cost one host pass over every arena.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <functional>
#include <numeric>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "build.h"
#include "device_common.h"
#include "general.h"
#include "internal.h"

namespace stormhip {

namespace {

void for_rows(uint32_t n, const std::function<bool(uint32_t, std::vector<uint64_t>&)>& fn,
              std::atomic<int>& bad_row, uint32_t nw) {
    unsigned th = std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
    if (n < 64) th = 1;
    std::atomic<uint32_t> next{0};
    auto work = [&] {
        std::vector<uint64_t> buf(nw ? nw : 1);
        for (;;) {
            const uint32_t lo = next.fetch_add(64);
            if (lo >= n || bad_row.load() >= 0) break;
            for (uint32_t r = lo; r < std::min(n, lo + 64); ++r)
                if (!fn(r, buf)) {
                    int expect = -1;
                    bad_row.compare_exchange_strong(expect, (int)r);
                    return;
                }
        }
    };
    std::vector<std::thread> ts;
    for (unsigned t = 1; t < th; ++t) ts.emplace_back(work);
    work();
    for (auto& t : ts) t.join();
}

bool offsets_ok(const uint64_t* off, uint32_t n) {
    for (uint32_t r = 0; r < n; ++r)
        if (off[r + 1] < off[r]) return false;
    return true;
}

uint64_t popcount_words(const uint64_t* w, uint32_t nw) {
    uint64_t c = 0;
    for (uint32_t k = 0; k < nw; ++k) c += (uint64_t)__builtin_popcountll(w[k]);
    return c;
}

// build_rank (kernels/storm_repr.cpp:25-45) of one bitmap into out.
void rank_of_words(const uint64_t* w, uint32_t nw, std::vector<uint64_t>& out) {
    const uint32_t nblk = (nw + 7) / 8;
    out.assign(2 * ((size_t)nblk + 1), 0);
    uint64_t acc = 0;
    for (uint32_t b = 0; b < nblk; ++b) {
        out[2 * (size_t)b] = acc;
        uint64_t sub = 0, rel = 0;
        for (uint32_t k = 0; k < 8; ++k) {
            if (k >= 1) sub |= (rel & 0x1FF) << (9 * (k - 1));
            if (b * 8 + k < nw) rel += (uint64_t)__builtin_popcountll(w[b * 8 + k]);
        }
        out[2 * (size_t)b + 1] = sub;
        acc += rel;
    }
    out[2 * (size_t)nblk] = acc;
}

// build_occ (kernels/storm_repr.cpp:47-59), 8-word bins.
void occ_of_words(const uint64_t* w, uint32_t nw, std::vector<uint64_t>& out) {
    const uint32_t bins = (nw + 7) / 8;
    out.assign(((size_t)bins + 63) / 64, 0);
    for (uint32_t b = 0; b < bins; ++b) {
        uint64_t any = 0;
        for (uint32_t k = b * 8; k < std::min(b * 8 + 8, nw); ++k) any |= w[k];
        if (any) out[b >> 6] |= 1ull << (b & 63);
    }
}

struct Checked {
    std::vector<uint64_t> mk_off;
    std::vector<uint32_t> mk_word, mk_pos;
};

}  // namespace

}  // namespace stormhip

using namespace stormhip;

extern "C" {

int storm_hip_corpus_upload(storm_hip_ctx* ctx, const storm_hip_arena* a, storm_hip_corpus** out) {
    if (!ctx || !a || !out) return STORM_HIP_ERR_INVALID;
    *out = nullptr;
    const uint32_t n = a->n_rows, U = a->universe;
    if (U == 0) {
        set_error("upload: universe must be positive");
        return STORM_HIP_ERR_INVALID;
    }
    if (!a->meta && n) return STORM_HIP_ERR_NULL_VALUES;
    const uint32_t nw = (U + 63) / 64;
    const bool lists = a->list_off, runs = a->run_off, ewah = a->ewah_off, comp = a->comp_off;
    const bool bms = a->bitmap_off, ranks = a->rank_off, occs = a->occ_off;
    // Offsets: non-decreasing, data present when any span is non-empty.
    struct Fam { const uint64_t* off; const void* data; const char* name; };
    const Fam fams[] = {{a->list_off, a->list, "list"}, {a->run_off, a->run_start, "run"},
                        {a->run_off, a->run_end, "run end"}, {a->ewah_off, a->ewah, "ewah"},
                        {a->comp_off, a->comp, "complement"}, {a->bitmap_off, a->bitmap, "bitmap"},
                        {a->rank_off, a->rank, "rank"}, {a->occ_off, a->occ, "occ"}};
    for (const Fam& f : fams) {
        if (!f.off) continue;
        if (!offsets_ok(f.off, n)) {
            set_error("upload: %s offsets decrease", f.name);
            return STORM_HIP_ERR_INVALID;
        }
        if (f.off[n] > f.off[0] && !f.data) {
            set_error("upload: %s arena is NULL with non-empty offsets", f.name);
            return STORM_HIP_ERR_NULL_VALUES;
        }
    }
    if (!lists && !runs && !bms && n) {
        set_error("upload: rows need a list, runs or a bitmap");
        return STORM_HIP_ERR_INVALID;
    }
    // Per row: rebuild its bitmap from the authoritative arena (list, else
    // runs, else its own bitmap) and check every other arena against it.
    Checked ck;
    std::vector<uint32_t> n_mk(n, 0);
    std::atomic<int> bad{-1};
    std::vector<std::string> err(1);
    std::atomic<int> err_set{0};
    auto fail_row = [&](uint32_t r, const char* what) {
        int z = 0;
        if (err_set.compare_exchange_strong(z, 1))
            err[0] = std::string("upload: row ") + std::to_string(r) + ": " + what;
        return false;
    };
    for_rows(n, [&](uint32_t r, std::vector<uint64_t>& w) -> bool {
        std::fill(w.begin(), w.end(), 0ull);
        const uint32_t* m = a->meta + (size_t)r * 8;
        const uint64_t bspan = bms ? a->bitmap_off[r + 1] - a->bitmap_off[r] : 0;
        if (bspan != 0 && bspan != nw) return fail_row(r, "bitmap span is not n_words words");
        const uint64_t* bw = bspan ? a->bitmap + a->bitmap_off[r] : nullptr;
        uint64_t card = 0, nruns = 0;
        uint32_t first = 0xFFFFFFFFu, last = 0;
        if (lists) {
            const uint32_t* p = a->list + a->list_off[r];
            const uint64_t c = a->list_off[r + 1] - a->list_off[r];
            for (uint64_t i = 0; i < c; ++i) {
                if (p[i] >= U || (i && p[i] <= p[i - 1])) return fail_row(r, "list not sorted, distinct, < universe");
                w[p[i] >> 6] |= 1ull << (p[i] & 63);
                if (i == 0 || p[i] != p[i - 1] + 1) ++nruns;
            }
            card = c;
            if (c) { first = p[0]; last = p[c - 1]; }
        }
        if (runs) {
            const uint64_t r0 = a->run_off[r], r1 = a->run_off[r + 1];
            uint64_t rc = 0;
            std::vector<uint64_t> tmp;
            if (lists) tmp.assign(nw, 0);
            std::vector<uint64_t>& dst = lists ? tmp : w;
            for (uint64_t k = r0; k < r1; ++k) {
                const uint32_t s = a->run_start[k], e = a->run_end[k];
                if (!(s < e && e <= U)) return fail_row(r, "run not inside the universe");
                if (k > r0 && s <= a->run_end[k - 1]) return fail_row(r, "runs not sorted and maximal");
                for (uint32_t x = s; x < e; ++x) dst[x >> 6] |= 1ull << (x & 63);
                rc += e - s;
            }
            if (lists) {
                if (tmp != w || r1 - r0 != nruns) return fail_row(r, "runs disagree with the list");
            } else {
                card = rc;
                nruns = r1 - r0;
                if (r1 > r0) { first = a->run_start[r0]; last = a->run_end[r1 - 1] - 1; }
            }
        }
        if (bw) {
            if (nw * 64 > U && (bw[nw - 1] >> (U & 63)) != 0) return fail_row(r, "bitmap bits past the universe");
            if (lists || runs) {
                if (std::memcmp(bw, w.data(), (size_t)nw * 8) != 0) return fail_row(r, "bitmap disagrees with the row");
            } else {
                std::memcpy(w.data(), bw, (size_t)nw * 8);
                card = popcount_words(bw, nw);
                uint64_t prev = 0;
                for (uint32_t q = 0; q < nw; ++q) {
                    nruns += (uint64_t)__builtin_popcountll(bw[q] & ~((bw[q] << 1) | prev));
                    prev = bw[q] >> 63;
                    if (bw[q]) {
                        if (first == 0xFFFFFFFFu) first = q * 64 + (uint32_t)__builtin_ctzll(bw[q]);
                        last = q * 64 + 63 - (uint32_t)__builtin_clzll(bw[q]);
                    }
                }
            }
        } else if (!lists && !runs) {
            return fail_row(r, "no list, runs or bitmap");
        }
        uint64_t nnz = 0;
        for (uint32_t q = 0; q < nw; ++q) nnz += w[q] != 0;
        // RowMeta (kernels/storm_repr.cpp:184-205).
        const uint32_t want_first = card ? first : 0xFFFFFFFFu, want_last = card ? last : 0;
        if (m[0] != card || m[1] != nruns || m[2] != nw || m[4] != want_first || m[5] != want_last)
            return fail_row(r, "meta disagrees with the row");
        const bool gate_rank = nruns && card / nruns >= 512, gate_occ = card * 512 > (uint64_t)U;
        // build_row's rows without a bitmap (sparse_only) carry nnz = n_runs
        // and no index flags; rows with one carry the exact live-word count
        // and the gates.
        const bool exact = m[3] == nnz && m[6] == (uint32_t)gate_rank && m[7] == (uint32_t)gate_occ;
        const bool sparse_only = !bw && m[3] == nruns && m[6] == 0 && m[7] == 0;
        if (!exact && !sparse_only) return fail_row(r, "meta index fields disagree with the row");
        if (ranks) {
            const uint64_t span = a->rank_off[r + 1] - a->rank_off[r];
            if (span) {
                if (!bw) return fail_row(r, "rank without a bitmap");
                std::vector<uint64_t> want;
                rank_of_words(w.data(), nw, want);
                if (span != want.size() || std::memcmp(a->rank + a->rank_off[r], want.data(), span * 8))
                    return fail_row(r, "rank9 disagrees with build_rank of the bitmap");
                if (!m[6]) return fail_row(r, "rank9 given for a row whose meta has no rank");
            } else if (bw && m[6]) {
                return fail_row(r, "meta has rank but no rank9 given");
            }
        }
        if (occs) {
            const uint64_t span = a->occ_off[r + 1] - a->occ_off[r];
            if (span) {
                if (!bw) return fail_row(r, "zone map without a bitmap");
                std::vector<uint64_t> want;
                occ_of_words(w.data(), nw, want);
                if (span != want.size() || std::memcmp(a->occ + a->occ_off[r], want.data(), span * 8))
                    return fail_row(r, "zone map disagrees with build_occ of the bitmap");
                if (!m[7]) return fail_row(r, "zone map given for a row whose meta has no occ");
            } else if (bw && m[7]) {
                return fail_row(r, "meta has occ but no zone map given");
            }
        }
        if (comp) {
            const uint32_t* p = a->comp + a->comp_off[r];
            const uint64_t c = a->comp_off[r + 1] - a->comp_off[r];
            const uint64_t want = 2 * card > U ? U - card : 0;
            if (c != want) return fail_row(r, "complement length");
            for (uint64_t i = 0; i < c; ++i)
                if (p[i] >= U || (i && p[i] <= p[i - 1]) || ((w[p[i] >> 6] >> (p[i] & 63)) & 1))
                    return fail_row(r, "complement is not the absent positions, sorted");
        }
        if (ewah) {
            // Decode (ewah_decode, kernels/storm_repr.cpp:61-74) and compare.
            const uint64_t* e = a->ewah + a->ewah_off[r];
            const uint64_t len = a->ewah_off[r + 1] - a->ewah_off[r];
            uint64_t o = 0, word = 0, mk = 0;
            while (o < len) {
                const uint64_t x = e[o++];
                const uint64_t fl = (x >> 1) & 0x7FFFFFFFull, ll = x >> 32;
                const uint64_t fv = (x & 1) ? ~0ull : 0ull;
                if (word + fl + ll > nw || o + ll > len) return fail_row(r, "EWAH overruns the row");
                for (uint64_t k = 0; k < fl; ++k)
                    if (w[word + k] != fv) return fail_row(r, "EWAH fill disagrees with the row");
                for (uint64_t k = 0; k < ll; ++k)
                    if (w[word + fl + k] != e[o + k]) return fail_row(r, "EWAH literal disagrees");
                o += ll;
                word += fl + ll;
                ++mk;
            }
            if (word != nw && !(nw == 0 && word == 0)) return fail_row(r, "EWAH decodes to the wrong length");
            if (len == 0) return fail_row(r, "EWAH buffer empty");
            n_mk[r] = (uint32_t)mk;
        }
        return true;
    }, bad, nw);
    if (bad.load() >= 0) {
        set_error("%s", err[0].c_str());
        return STORM_HIP_ERR_INVALID;
    }
    if (ewah) {
        ck.mk_off.assign((size_t)n + 1, 0);
        for (uint32_t r = 0; r < n; ++r) ck.mk_off[r + 1] = ck.mk_off[r] + n_mk[r];
        ck.mk_word.resize(ck.mk_off[n]);
        ck.mk_pos.resize(ck.mk_off[n]);
        for (uint32_t r = 0; r < n; ++r) {
            const uint64_t* e = a->ewah + a->ewah_off[r];
            const uint64_t len = a->ewah_off[r + 1] - a->ewah_off[r];
            uint64_t o = 0, word = 0, k = ck.mk_off[r];
            while (o < len) {
                const uint64_t x = e[o];
                ck.mk_word[k] = (uint32_t)word;
                ck.mk_pos[k] = (uint32_t)o;
                ++k;
                o += 1 + (x >> 32);
                word += ((x >> 1) & 0x7FFFFFFFull) + (x >> 32);
            }
        }
    }
    SH_HIP(hipSetDevice(ctx->device));
    uint32_t flags = STORM_HIP_BITMAP_NONE;
    if (runs) flags |= STORM_HIP_BUILD_RUNS;
    if (ewah) flags |= STORM_HIP_BUILD_EWAH;
    if (comp) flags |= STORM_HIP_BUILD_COMP;
    storm_hip_corpus* c = corpus_new(ctx, n, U, flags);
    if (!c) return STORM_HIP_ERR_NOMEM;
    c->has_lists = lists;
    hipStream_t s = ctx->stream;
    uint64_t& B = c->device_bytes;
    auto fail = [&](int st) { storm_hip_corpus_free(c); return st; };
    // Offsets rebased to 0 and the arenas' used ranges.
    auto up_off = [&](uint64_t*& d, const uint64_t* off) -> int {
        std::vector<uint64_t> v((size_t)n + 1, 0);
        if (off) for (uint32_t r = 0; r <= n; ++r) v[r] = off[r] - off[0];
        return upload_vec(d, v, s, B);
    };
    auto up_span = [&](auto*& d, const auto* src, const uint64_t* off) -> int {
        using T = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
        const uint64_t cnt = off ? off[n] - off[0] : 0;
        std::vector<T> v;
        if (cnt) v.assign(src + off[0], src + off[n]);
        return upload_vec(d, v, s, B);
    };
    int st = STORM_HIP_OK;
    if (st == STORM_HIP_OK) st = up_off(c->d_list_off, lists ? a->list_off : nullptr);
    if (st == STORM_HIP_OK) st = up_span(c->d_list, a->list, lists ? a->list_off : nullptr);
    if (st == STORM_HIP_OK) st = up_off(c->d_run_off, runs ? a->run_off : nullptr);
    if (st == STORM_HIP_OK) st = up_span(c->d_run_start, a->run_start, runs ? a->run_off : nullptr);
    if (st == STORM_HIP_OK) st = up_span(c->d_run_end, a->run_end, runs ? a->run_off : nullptr);
    if (st == STORM_HIP_OK) st = up_off(c->d_ewah_off, ewah ? a->ewah_off : nullptr);
    if (st == STORM_HIP_OK) st = up_span(c->d_ewah, a->ewah, ewah ? a->ewah_off : nullptr);
    if (st == STORM_HIP_OK) st = upload_vec(c->d_mk_off, ewah ? ck.mk_off : std::vector<uint64_t>((size_t)n + 1, 0), s, B);
    if (st == STORM_HIP_OK) st = upload_vec(c->d_mk_word, ck.mk_word, s, B);
    if (st == STORM_HIP_OK) st = upload_vec(c->d_mk_pos, ck.mk_pos, s, B);
    if (st == STORM_HIP_OK) st = up_off(c->d_comp_off, comp ? a->comp_off : nullptr);
    if (st == STORM_HIP_OK) st = up_span(c->d_comp, a->comp, comp ? a->comp_off : nullptr);
    if (st == STORM_HIP_OK) {
        c->h_meta.assign(a->meta, a->meta + (size_t)n * 8);
        st = upload_vec(c->d_meta, c->h_meta, s, B);
    }
    if (st != STORM_HIP_OK) return fail(st);
    std::vector<uint8_t> want(n, 0);
    for (uint32_t r = 0; r < n; ++r) want[r] = bms && a->bitmap_off[r + 1] > a->bitmap_off[r];
    if (lists || runs) {
        // Bitmap rows are filled from the lists or runs (checked equal).
        st = corpus_finish(ctx, c, want);
    } else {
        // Bitmaps only (every row has one): slot = row, copied as given.
        c->h_bm_row.assign((size_t)n + 1, 0);
        std::iota(c->h_bm_row.begin(), c->h_bm_row.end(), 0u);
        c->n_bitmap_rows = n;
        c->bm_capacity = n;
        c->bitmaps_complete = true;
        st = upload_vec(c->d_bm_row, c->h_bm_row, s, B);
        const uint64_t bytes = ((uint64_t)n + 1) * c->ld * 8;
        if (st == STORM_HIP_OK && hipMalloc((void**)&c->d_bitmap, bytes) != hipSuccess)
            st = STORM_HIP_ERR_NOMEM;
        if (st == STORM_HIP_OK) {
            B += bytes;
            if (hipMemsetAsync(c->d_bitmap, 0, bytes, s) != hipSuccess) st = STORM_HIP_ERR_DEVICE;
            for (uint32_t r = 0; r < n && st == STORM_HIP_OK; ++r)
                if (hipMemcpyAsync(c->d_bitmap + (uint64_t)r * c->ld, a->bitmap + a->bitmap_off[r],
                                   (size_t)nw * 8, hipMemcpyHostToDevice, s) != hipSuccess)
                    st = STORM_HIP_ERR_DEVICE;
        }
        if (st == STORM_HIP_OK) st = corpus_finish_store(ctx, c);
    }
    if (st != STORM_HIP_OK) return fail(st);
    *out = c;
    return STORM_HIP_OK;
}

}  // extern "C"
