"""Representations on the GPU beyond lists: rows held only as runs or only as
bitmaps, the arena upload fed the reference's own build_row arenas, rank9 and
zone maps, the B x R rank path, the C x B gap path, the device generators and
the column closed forms -- each against the CPU oracle or the compiled
reference."""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference checker not built")

CELLS = ["BB", "BS", "BR", "BW", "SS", "SR", "SW", "RR", "RW", "WW", "EMPTY", "CB"]
LIST_CELLS = {"BS", "SS", "SR", "SW"}


def _case(*a):
    print("case", *a, flush=True)


def _positions_from_runs(off, rs, re_):
    n = len(off) - 1
    rows = []
    for r in range(n):
        a, b = int(off[r]), int(off[r + 1])
        rows.append(np.concatenate([np.arange(s, e, dtype=np.uint32)
                                    for s, e in zip(rs[a:b], re_[a:b])])
                    if b > a else np.zeros(0, np.uint32))
    return rows


def _csr(rows):
    offs = np.zeros(len(rows) + 1, np.uint64)
    offs[1:] = np.cumsum([len(r) for r in rows])
    pos = np.concatenate(rows).astype(np.uint32) if offs[-1] else np.zeros(0, np.uint32)
    return offs, pos


def _rows_of(offs, pos):
    return [pos[int(offs[r]):int(offs[r + 1])] for r in range(len(offs) - 1)]


def _check_allpairs(c, offs, pos, u, policies=(("analytic", {}),), tiles=(64, 256)):
    want_sum, want = O.allpairs_lists(offs, pos)
    total, rs = O.lists_closed_forms(offs, pos, u)
    assert want_sum == total
    got_total, got_rs = c.closed_form()
    assert got_total == total
    assert np.array_equal(got_rs, rs)
    n = len(offs) - 1
    for policy, kw in policies:
        for tile in tiles:
            st = c.allpairs(policy=policy, out="tri", tile=tile, **kw)
            assert st["sum"] == want_sum, (policy, kw, tile)
            if n >= 2:
                assert np.array_equal(c.tri(), want), (policy, kw, tile)
                assert np.array_equal(c.row_sums(), rs), (policy, kw, tile)
            yield st


# ---- device generators -------------------------------------------------------
RUNS_CASES = [  # (n, universe, lambda, mean run, singleton fraction, seed)
    (50, 100003, 5.0, 40.0, 0.05, 11),
    (40, 65536, 45.0, 300.0, 0.1, 12),      # lambda > 30: sum of Poisson draws
    (30, 1000, 2.0, 1.0, 0.5, 13),          # unit runs
    (64, 4096, 0.5, 2000.0, 0.0, 14),       # few long runs, clipped at the end
    (3, 70000, 0.0, 10.0, 0.05, 15),        # empty rows
]
SKIP_CASES = [(100, 50001, 0.001, 21), (64, 128, 0.03, 22), (3, 70000, 0.0, 23),
              (70, 1 << 16, 0.0005, 24)]
BITS_CASES = [(70, 5000, 0.05, 31), (65, 64 * 20, 0.5, 32), (10, 100, 1.0, 33),
              (129, 4033, 1 / 32, 34)]


@pytest.mark.parametrize("case", RUNS_CASES)
def test_device_runs_generator_matches_host(sh, ctx, case):
    n, u, lam, L, frac, seed = case
    _case("gen runs", case)
    offs, pos = sh.gen_corpus("runs", n, u, lam, L, frac, seed)
    with ctx.gen_device("runs", n, u, lam, L, frac, seed, bitmap_policy=0) as c:
        a = c.arenas()
        rows = _positions_from_runs(a["run_off"], a["run_start"], a["run_end"])
        for r in range(n):
            assert np.array_equal(rows[r], pos[int(offs[r]):int(offs[r + 1])]), (case, r)
        assert len(a["list"]) == 0
        # runs are maximal
        for r in range(n):
            s, e = a["run_start"][a["run_off"][r]:a["run_off"][r + 1]], \
                a["run_end"][a["run_off"][r]:a["run_off"][r + 1]]
            assert np.all(s[1:] > e[:-1])
        for r in range(n):
            m = O.build_row(rows[r], u)
            want = np.concatenate([m["meta"], [m["has_rank"], m["has_occ"]]]).astype(np.uint32)
            assert np.array_equal(a["meta"][8 * r:8 * r + 8], want), (case, r)
        for st in _check_allpairs(c, offs, pos, u):
            used = set(st["cell_tiles"]) - {"EMPTY"}
            assert not (used & LIST_CELLS), st["cell_tiles"]


@pytest.mark.parametrize("case", SKIP_CASES)
def test_device_sparse_bernoulli_matches_host(sh, ctx, case):
    n, u, d, seed = case
    _case("gen skip", case)
    offs, pos = sh.gen_corpus("bernoulli", n, u, d, 0.0, 0.0, seed)
    with ctx.gen_device("bernoulli", n, u, d, seed=seed, bitmap_policy=0 | 4 | 8 | 16) as c:
        a = c.arenas()
        assert np.array_equal(a["list_off"], offs)
        assert np.array_equal(a["list"], pos)
        list(_check_allpairs(c, offs, pos, u))


@pytest.mark.parametrize("case", BITS_CASES)
def test_device_dense_bernoulli_matches_stream(ctx, case):
    n, u, p, seed = case
    _case("gen bits", case)
    words = O.bernoulli_rows(seed, p, u, 0, n)
    rows = [np.flatnonzero(np.unpackbits(w.view(np.uint8), bitorder="little")[:u]).astype(np.uint32)
            for w in words]
    offs, pos = _csr(rows)
    with ctx.gen_device("bernoulli", n, u, p, seed=seed) as c:
        a = c.arenas()
        nw = (u + 63) // 64
        ld = len(a["bitmap"]) // max(n, 1)
        # The store is laid out in density order: row r in slot bm_row[r],
        # which is its position in the corpus's order.
        assert np.array_equal(np.sort(a["bm_row"]), np.arange(n, dtype=np.uint32))
        assert np.array_equal(a["bm_row"][c.order()], np.arange(n, dtype=np.uint32))
        got = a["bitmap"].reshape(n, ld)[a["bm_row"].astype(np.int64), :nw]
        assert np.array_equal(got, words)
        for r in range(n):
            m = O.build_row(rows[r], u)
            want = np.concatenate([m["meta"], [m["has_rank"], m["has_occ"]]]).astype(np.uint32)
            assert np.array_equal(a["meta"][8 * r:8 * r + 8], want), (case, r)
        for st in _check_allpairs(c, offs, pos, u):
            assert set(st["cell_tiles"]) - {"EMPTY"} <= {"BB"}, st["cell_tiles"]


# ---- rank9, zone maps, B x R rank path, C x B gap path ---------------------------
def test_rank_and_occ_and_br_rank_path(sh, ctx):
    """Long runs (mean run length >= 512): every row's bitmap carries rank9 and
    a zone map equal to the oracle's build_rank / build_occ, and B x R forced
    on every tile (rank lookups for runs of 256 bits or more) is exact."""
    n, u = 120, 300007
    offs, pos = sh.gen_corpus("runs", n, u, 4.0, 3000.0, 0.0, 41)
    L = O.oracle()
    with ctx.gen_device("runs", n, u, 4.0, 3000.0, 0.0, 41, bitmap_policy=1) as c:
        a = c.arenas()
        nw = (u + 63) // 64
        ld = len(a["bitmap"]) // n
        nblk = (nw + 7) // 8
        rank_ld = 2 * (nblk + 1)
        occ_ld = (nblk + 63) // 64
        assert (a["meta"].reshape(n, 8)[:, 6] > 0).sum() > n // 2   # most rows ranked
        for r in range(n):
            bm = np.ascontiguousarray(a["bitmap"][a["bm_row"][r] * ld:][:nw])
            m = a["meta"][8 * r:8 * r + 8]
            if m[6]:
                want = np.zeros(rank_ld, np.uint64)
                L.orc_build_rank(O._p(bm), nw, O._p(want))
                k = a["rank_row"][r]
                assert np.array_equal(a["rank"][k * rank_ld:(k + 1) * rank_ld], want), r
            else:
                assert a["rank_row"][r] == 0xFFFFFFFF
            if m[7]:
                want = np.zeros(occ_ld, np.uint64)
                L.orc_build_occ(O._p(bm), nw, 8, O._p(want))
                k = a["occ_row"][r]
                assert np.array_equal(a["occ"][k * occ_ld:(k + 1) * occ_ld], want), r
        for st in _check_allpairs(c, offs, pos, u, policies=[("fixed", {"fixed_cell": 2}),
                                                              ("analytic", {})]):
            pass
        st = c.allpairs(policy="fixed", fixed_cell=2, out="sum", tile=64)
        assert st["cell_tiles"].get("BR", 0) > 0


def test_cb_gap_path_without_complement_lists(sh, ctx):
    """Rows above half density held as runs with bitmaps and no complement
    list: C x B walks the gaps between the runs (rank lookups on long gaps)."""
    n, u = 80, 20011
    offs, pos = sh.gen_corpus("runs", n, u, 8.0, 3000.0, 0.0, 43)
    dense = sum(1 for r in _rows_of(offs, pos) if 2 * len(r) > u)
    assert dense > n // 2
    with ctx.gen_device("runs", n, u, 8.0, 3000.0, 0.0, 43, bitmap_policy=1) as c:
        for st in _check_allpairs(c, offs, pos, u, policies=[("fixed", {"fixed_cell": 11})],
                                  tiles=(64,)):
            assert st["cell_tiles"].get("CB", 0) > 0


# ---- arena upload ------------------------------------------------------------------
def _ref_corpus():
    rows = []
    for u in (4096,):
        rows += [r for r in (list(range(0, u, 3)), list(range(100, 4000)),
                             list(range(0, u)), [], [5], list(range(64, 128)),
                             list(range(0, 2048)) + list(range(2600, 4096)))]
    offs, pos = O.ref_generate(60, 4096, 0.05, 2, 1, 0x5151)
    rows += [list(x) for x in _rows_of(offs, pos)]
    # long runs (rank gate) and rows above half density (complements)
    rng = np.random.default_rng(3)
    for _ in range(20):
        s = int(rng.integers(0, 1000))
        rows.append(list(range(s, s + int(rng.integers(1500, 3000)))))
    return 4096, [np.array(sorted(set(r)), np.uint32) for r in rows]


def _ref_arenas(u, rows, families, bitmap_rows=None, sparse_only=()):
    """Arenas of the rows as the reference's build_row writes them."""
    refs = [O.ref_build_row(r, u) for r in rows]
    n = len(rows)
    out = {}

    def fam(key, parts, dt):
        off = np.zeros(n + 1, np.uint64)
        off[1:] = np.cumsum([len(p) for p in parts])
        out[key + "_off"] = off
        out[key] = np.concatenate(parts).astype(dt) if off[-1] else np.zeros(0, dt)

    has_bm = [bitmap_rows is None or r in bitmap_rows for r in range(n)]
    for r in sparse_only:
        has_bm[r] = False
    if "list" in families:
        fam("list", rows, np.uint32)
    if "runs" in families:
        fam("run", [x["run_start"] for x in refs], np.uint32)
        out["run_start"] = out.pop("run")
        out["run_end"] = np.concatenate([x["run_end"] for x in refs]).astype(np.uint32) \
            if out["run_off"][-1] else np.zeros(0, np.uint32)
    if "ewah" in families:
        fam("ewah", [x["ewah"] for x in refs], np.uint64)
    if "comp" in families:
        fam("comp", [x["comp"] for x in refs], np.uint32)
    if "bitmap" in families:
        fam("bitmap", [x["bitmap"] if has_bm[r] else np.zeros(0, np.uint64)
                       for r, x in enumerate(refs)], np.uint64)
        fam("rank", [x["rank"] if has_bm[r] else np.zeros(0, np.uint64)
                     for r, x in enumerate(refs)], np.uint64)
        fam("occ", [x["occ"] if has_bm[r] else np.zeros(0, np.uint64)
                    for r, x in enumerate(refs)], np.uint64)
    meta = np.zeros((n, 8), np.uint32)
    for r, x in enumerate(refs):
        meta[r, :6] = x["meta"]
        meta[r, 6:] = x["sizes"][7:9]
        if not has_bm[r]:
            # build_row(sparse_only): no bitmap, nnz = n_runs, no indexes.
            meta[r, 3] = x["meta"][1]
            meta[r, 6:] = 0
    out["meta"] = meta.reshape(-1)
    return out, refs


@needs_ref
@pytest.mark.parametrize("families", [
    ("list", "runs", "ewah", "comp", "bitmap"),
    ("runs", "bitmap"),
    ("bitmap",),
    ("runs",),
    ("list", "runs", "ewah", "comp"),
])
def test_upload_reference_arenas(ctx, families):
    """storm_hip_corpus_upload fed the reference's build_row arenas: every
    arena comes back byte for byte (rank9 and zone maps rebuilt on the device
    equal the reference's), and every policy is exact; without lists the
    list cells are never used."""
    u, rows = _ref_corpus()
    n = len(rows)
    _case("upload", families)
    arenas, refs = _ref_arenas(u, rows, families)
    offs, pos = _csr(rows)
    with ctx.upload_arenas(u, n, arenas) as c:
        a = c.arenas()
        if "list" in families:
            assert np.array_equal(a["list_off"], arenas["list_off"])
            assert np.array_equal(a["list"], arenas["list"])
        if "runs" in families:
            assert np.array_equal(a["run_start"], arenas["run_start"])
            assert np.array_equal(a["run_end"], arenas["run_end"])
        if "ewah" in families:
            assert np.array_equal(a["ewah"], arenas["ewah"])
        if "comp" in families:
            assert np.array_equal(a["comp"], arenas["comp"])
        assert np.array_equal(a["meta"], arenas["meta"])
        if "bitmap" in families:
            nw = (u + 63) // 64
            ld = len(a["bitmap"]) // max(1, len(set(a["bm_row"]) - {0xFFFFFFFF}))
            nblk = (nw + 7) // 8
            for r, x in enumerate(refs):
                k = a["bm_row"][r]
                assert np.array_equal(a["bitmap"][k * ld:k * ld + nw], x["bitmap"]), r
                if len(x["rank"]):
                    k = a["rank_row"][r]
                    assert np.array_equal(a["rank"][k * 2 * (nblk + 1):(k + 1) * 2 * (nblk + 1)],
                                          x["rank"]), r
                if len(x["occ"]):
                    k = a["occ_row"][r]
                    lo = len(x["occ"])
                    assert np.array_equal(a["occ"][k * lo:(k + 1) * lo], x["occ"]), r
        pol = [("analytic", {})] + [("fixed", {"fixed_cell": k}) for k in range(12) if k != 10]
        for st in _check_allpairs(c, offs, pos, u, policies=pol, tiles=(64,)):
            if "list" not in families:
                assert not (set(st["cell_tiles"]) & LIST_CELLS), st["cell_tiles"]


@needs_ref
def test_upload_sparse_only_rows(ctx):
    """Rows without a bitmap carry build_row's sparse-only meta (n_nonzero_w =
    n_runs, no indexes); they are accepted and counted exactly."""
    u, rows = _ref_corpus()
    arenas, _ = _ref_arenas(u, rows, ("list", "runs", "bitmap"), sparse_only=range(0, 40, 3))
    offs, pos = _csr(rows)
    with ctx.upload_arenas(u, len(rows), arenas) as c:
        list(_check_allpairs(c, offs, pos, u, tiles=(64,)))


@needs_ref
def test_upload_rejects_bad_arenas(sh, ctx):
    u, rows = _ref_corpus()
    good, _ = _ref_arenas(u, rows, ("list", "runs", "bitmap"))
    n = len(rows)

    def bad(mutate, code):
        a = {k: (v.copy() if v is not None else None) for k, v in good.items()}
        mutate(a)
        with pytest.raises(sh.StormHipError) as e:
            ctx.upload_arenas(u, n, a)
        assert f"({code})" in str(e.value), str(e.value)

    def dec_offsets(a):
        a["list_off"][5] = a["list_off"][6] + 1
    bad(dec_offsets, -1)

    def null_list(a):
        a["list"] = None
    bad(null_list, -2)

    def card_off(a):
        a["meta"][8 * 3] += 1
    bad(card_off, -1)

    def unsorted(a):
        lo = int(a["list_off"][1])
        a["list"][lo], a["list"][lo + 1] = a["list"][lo + 1], a["list"][lo]
    bad(unsorted, -1)

    def run_overlap(a):
        # row 0 is every third position: runs [0,1), [3,4), ...; the second
        # run made to start inside the first one's end.
        k = int(a["run_off"][0])
        a["run_start"][k + 1] = a["run_end"][k]
    bad(run_overlap, -1)

    def rank_wrong(a):
        k = int(np.flatnonzero(np.diff(a["rank_off"]))[0])
        a["rank"][int(a["rank_off"][k]) + 2] += 1
    bad(rank_wrong, -1)

    def past_universe(a):
        ub = (u + 63) // 64 * 64
        if ub > u:
            a["bitmap"][int(a["bitmap_off"][1]) - 1] |= np.uint64(1) << np.uint64(63)
        else:
            a["meta"][8 * 0 + 2] += 1
    bad(past_universe, -1)


def test_wide_bb_tiles_in_k_slices(sh, ctx):
    """256-row B x B tiles of wide rows (78,125 words, 3 K slices per tile)
    queued as slices that add into the zeroed triangle: exact pair by pair,
    alone and in a batch with a second corpus whose rows are narrow."""
    n, u, p, seed = 300, 5_000_000, 0.04, 51
    words = O.bernoulli_rows(seed, p, u, 0, n)
    want_sum, want = O.allpairs_words(words)
    total, rs = O.column_closed_forms(words, threads=8)
    with ctx.gen_device("bernoulli", n, u, p, seed=seed) as c, \
            ctx.gen_device("bernoulli", 200, 4096, 0.5, seed=52) as c2:
        for policy in ("analytic", "fixed"):
            st = c.allpairs(policy=policy, out="tri", tile=256, fixed_cell=0)
            assert st["sum"] == want_sum == total
            assert np.array_equal(c.tri(), want)
            assert np.array_equal(c.row_sums(), rs)
        w2 = O.bernoulli_rows(52, 0.5, 4096, 0, 200)
        s2, t2 = O.allpairs_words(w2)
        with ctx.batch([c, c2]) as b:
            for _ in range(2):   # graph replays re-zero the sliced triangle
                st, sums, lat = b.allpairs(policy="analytic", out="tri", tile=256, use_graph=True)
                assert [int(x) for x in sums] == [want_sum, s2]
                assert np.array_equal(c.tri(), want)
                assert np.array_equal(c2.tri(), t2)


def test_wide_corpus_bb_fp4_k_slices_and_int8_queue(sh, ctx):
    """Rows of 2^18 words or more (whole-row counts may pass fp4's exact f32
    range): their 256-row B x B tiles run as K slices, each below 2^18 words,
    which stay exact in fp4 and add into the output as integers -- one fp4
    queue beside a narrow corpus of the same batch. Rows past 64 slices of
    2^18 words (here 2^30 + 12,345 bits) keep the int8 queue beside the fp4
    one. All exact."""
    n, u, p, seed = 100, (1 << 24) + 4097, 0.04, 53   # p >= 1/32: rows held as bitmaps
    words = O.bernoulli_rows(seed, p, u, 0, n)
    want_sum, want = O.allpairs_words(words)
    w2 = O.bernoulli_rows(54, 0.5, 4096, 0, 300)
    s2, t2 = O.allpairs_words(w2)
    with ctx.gen_device("bernoulli", n, u, p, seed=seed) as c, \
            ctx.gen_device("bernoulli", 300, 4096, 0.5, seed=54) as c2:
        st = c.allpairs(policy="analytic", out="tri", tile=256)
        assert st["sum"] == want_sum
        assert np.array_equal(c.tri(), want)
        with ctx.batch([c, c2]) as b:
            for graph in (False, True, True):
                st, sums, lat = b.allpairs(policy="analytic", out="tri", tile=256,
                                           use_graph=graph)
                assert [int(x) for x in sums] == [want_sum, s2]
                assert np.array_equal(c.tri(), want)
                assert np.array_equal(c2.tri(), t2)
                assert st["bb_launches"] == 1
        # Rows of 2^30 + 12,345 bits: 64 K slices of more than 2^18 words.
        rng = np.random.default_rng(55)
        uw = (1 << 30) + 12_345
        pool = np.unique(rng.integers(0, uw, size=60_000)).astype(np.uint32)
        rows = [np.sort(rng.choice(pool, size=20_000, replace=False)).astype(np.uint32)
                for _ in range(6)]
        offs = np.zeros(7, np.uint64)
        offs[1:] = np.cumsum([len(r) for r in rows])
        pos = np.concatenate(rows)
        ws, wt = O.allpairs_lists(offs, pos)
        with ctx.from_lists(offs, pos, uw, bitmap_policy=1 | 4 | 8 | 16) as c3:
            with ctx.batch([c3, c2]) as b:
                st, sums, lat = b.allpairs(policy="fixed", fixed_cell=0, out="tri", tile=256)
                assert [int(x) for x in sums] == [ws, s2]
                assert np.array_equal(c3.tri(), wt)
                assert np.array_equal(c2.tri(), t2)
                assert st["bb_launches"] == 2
