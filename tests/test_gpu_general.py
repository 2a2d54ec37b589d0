"""Mixed-representation corpora on the GPU: every cell, every selecting policy,
every tile width, against the CPU oracle's per-pair counts."""
import ctypes as C

import numpy as np
import pytest

import corpora
import oracle_py as O

pytestmark = pytest.mark.gpu

CELLS = ["BB", "BS", "BR", "BW", "SS", "SR", "SW", "RR", "RW", "WW", "EMPTY", "CB"]
BUILD_ALL = 4 | 8 | 16
CORPORA = corpora.all_corpora()


def _oracle(rows):
    offs, pos = corpora.csr(rows)
    s, tri = O.allpairs_lists(offs, pos)
    return offs, pos, s, tri


@pytest.fixture(scope="module")
def cases():
    return [(name, u) + _oracle(rows) for name, rows, u in CORPORA]


@pytest.mark.parametrize("bitmaps", [0, 1, 2])
def test_analytic_policy_every_tile_width(ctx, cases, bitmaps):
    for name, u, offs, pos, want_sum, want in cases:
        print("case analytic", name, "bitmaps", bitmaps, flush=True)
        with ctx.from_lists(offs, pos, u, bitmap_policy=bitmaps | BUILD_ALL) as c:
            for tile in (64, 128, 256):
                st = c.allpairs(policy="analytic", out="tri", tile=tile)
                assert st["sum"] == want_sum, (name, tile)
                assert np.array_equal(c.tri(), want), (name, tile)
                assert sum(st["cell_pairs"].values()) == st["pairs"], (name, tile, st)


def test_fixed_policy_every_cell(ctx, cases):
    """Every cell forced on every tile it is eligible for (others fall back)."""
    for name, u, offs, pos, want_sum, want in cases:
        print("case fixed", name, flush=True)
        with ctx.from_lists(offs, pos, u, bitmap_policy=1 | BUILD_ALL) as c:
            for k, cell in enumerate(CELLS):
                if cell == "EMPTY":
                    continue
                st = c.allpairs(policy="fixed", fixed_cell=k, out="tri", tile=64)
                assert st["sum"] == want_sum, (name, cell)
                assert np.array_equal(c.tri(), want), (name, cell)
                if cell != "CB":
                    # with every representation resident, only tiles proven
                    # empty avoid the forced cell
                    used = set(st["cell_tiles"]) - {"EMPTY"}
                    assert used <= {cell}, (name, cell, st["cell_tiles"])


def test_all_bitmap_policy_materialises(ctx, cases):
    for name, u, offs, pos, want_sum, want in cases[:12]:
        print("case allbitmap", name, flush=True)
        with ctx.from_lists(offs, pos, u, bitmap_policy=2) as c:
            for bb in ("mfma", "valu"):
                st = c.allpairs(policy="all_bitmap", out="tri", bb=bb)
                assert st["sum"] == want_sum
                assert np.array_equal(c.tri(), want)
                assert st["skipped"] == 0


def test_analytic_decisions_match_oracle_cost_model(ctx, cases):
    """With every representation resident the device decision is the
    reference's select_pairing on tile_meta (storm_allpairs.cpp:552-596), except
    that a tile the span test proves empty is settled without a kernel."""
    L = O.oracle()
    model = (C.c_double * 16)()
    L.orc_default_model(model)

    class M(C.Structure):
        _fields_ = [("a", C.c_uint32 * 6), ("hr", C.c_uint8), ("ho", C.c_uint8)]

    for name, u, offs, pos, _, _ in cases:
        rows = [pos[int(offs[i]):int(offs[i + 1])] for i in range(len(offs) - 1)]
        n = len(rows)
        metas = (M * n)()
        for i, r in enumerate(rows):
            nw = (u + 63) // 64
            k = len(r)
            runs = int(np.count_nonzero(np.diff(r.astype(np.int64)) != 1) + 1) if k else 0
            metas[i].a[0], metas[i].a[1], metas[i].a[2] = k, runs, nw
            metas[i].a[3] = len(np.unique(r >> 6))
            metas[i].a[4] = int(r[0]) if k else 0xFFFFFFFF
            metas[i].a[5] = int(r[-1]) if k else 0
            metas[i].hr = int(runs and k // runs >= 512)
            metas[i].ho = int(k * 512 > u)
        order = np.zeros(n, np.uint32)
        L.orc_density_order(metas, n, order.ctypes.data_as(C.c_void_p))
        for tile in (64, 128):
            with ctx.from_lists(offs, pos, u, bitmap_policy=1 | BUILD_ALL) as c:
                c.allpairs(policy="analytic", tile=tile)
                T = (n + tile - 1) // tile
                got = c.last_decisions(T * (T + 1) // 2)
            t = 0
            for bi in range(T):
                for bj in range(bi, T):
                    ta, tb = M(), M()
                    lo_a, hi_a = bi * tile, min(n, bi * tile + tile)
                    lo_b, hi_b = bj * tile, min(n, bj * tile + tile)
                    L.orc_tile_meta(metas, order.ctypes.data_as(C.c_void_p), lo_a, hi_a, C.byref(ta))
                    L.orc_tile_meta(metas, order.ctypes.data_as(C.c_void_p), lo_b, hi_b, C.byref(tb))
                    want = L.orc_select_pairing(model, C.byref(ta), C.byref(tb), None)
                    rows_a = [rows[order[k]] for k in range(lo_a, hi_a)]
                    rows_b = [rows[order[k]] for k in range(lo_b, hi_b)]
                    ka = [r for r in rows_a if len(r)]
                    kb = [r for r in rows_b if len(r)]
                    sound_empty = (not ka or not kb or
                                   max(int(r[-1]) for r in ka) < min(int(r[0]) for r in kb) or
                                   max(int(r[-1]) for r in kb) < min(int(r[0]) for r in ka))
                    if sound_empty:
                        want = 10
                    elif want == 10:
                        want = 0      # the reference runs B x B on such a tile
                    assert got[t] == want, (name, tile, bi, bj)
                    t += 1


def test_graph_replay_matches_eager(ctx, cases):
    name, u, offs, pos, want_sum, want = cases[20]
    with ctx.from_lists(offs, pos, u, bitmap_policy=0 | BUILD_ALL) as c:
        for rep in range(4):
            st = c.allpairs(policy="analytic", out="tri", use_graph=True)
            assert st["sum"] == want_sum, rep
            assert np.array_equal(c.tri(), want), rep
            st2 = c.allpairs(policy="fixed", fixed_cell=4, out="tri")
            assert st2["sum"] == want_sum
            ctx.sync()


@pytest.mark.parametrize("tile", [64, 256])
def test_rr_join_and_staged(ctx, tile):
    """R x R over 320 run rows whose cardinalities span 1..30,000 over a
    universe of 200,003: tile pairs of similar rows take the interval join,
    pairs of a tiny-row tile with a large-row tile the staged search; long
    runs crossing chunk and bucket edges, single-position runs, empty rows,
    and rows covering the whole universe are all present."""
    rng = np.random.default_rng(77)
    u = 200_003
    rows = []
    for r in range(320):
        k = int(np.exp(rng.random() * np.log(30_000)))
        mean_run = (1, 4, 64, 2_000)[r % 4]
        rows.append(corpora.runs_row(rng, u, k, mean_run))
    rows[5] = np.zeros(0, np.uint32)
    rows[17] = np.arange(u, dtype=np.uint32)
    rows[29] = np.arange(u - 70_000, u, dtype=np.uint32)
    offs, pos = corpora.csr(rows)
    want_sum, want = O.allpairs_lists(offs, pos)
    print("case rr join tile", tile, flush=True)
    with ctx.from_lists(offs, pos, u, bitmap_policy=0 | BUILD_ALL) as c:
        st = c.allpairs(policy="fixed", fixed_cell=CELLS.index("RR"), out="tri", tile=tile)
        assert st["sum"] == want_sum
        assert np.array_equal(c.tri(), want)
        assert set(st["cell_tiles"]) - {"EMPTY"} <= {"RR"}
        st = c.allpairs(policy="analytic", out="tri", tile=tile)
        assert st["sum"] == want_sum
        assert np.array_equal(c.tri(), want)


def test_tiny_and_empty_corpora(ctx):
    for rows, u in [([], 64), ([np.array([1], np.uint32)], 64),
                    ([np.zeros(0, np.uint32)] * 5, 100),
                    ([np.array([3], np.uint32), np.array([3], np.uint32)], 64)]:
        offs, pos = corpora.csr(rows)
        with ctx.from_lists(offs, pos, u, bitmap_policy=0 | BUILD_ALL) as c:
            want = 1 if len(rows) == 2 else 0
            for pol in ("analytic", "all_bitmap"):
                st = c.allpairs(policy=pol, out="tri")
                assert st["sum"] == want
                if len(rows) >= 2:
                    assert int(c.tri().sum()) == want


def test_upload_rejects_bad_rows(sh, ctx):
    offs = np.array([0, 2], np.uint64)
    with pytest.raises(sh.StormHipError):
        ctx.from_lists(offs, np.array([5, 5], np.uint32), 64)     # not distinct
    with pytest.raises(sh.StormHipError):
        ctx.from_lists(offs, np.array([1, 64], np.uint32), 64)    # out of range


@pytest.mark.parametrize("tile", [64, 256])
def test_batch_matches_single_calls(ctx, cases, tile):
    """Many corpora (different universes and representations) through one set
    of launches: each corpus's sum and triangle equal the oracle's, the stats
    add up, and graph replays agree with eager runs."""
    picks = [cases[k] for k in (0, 3, 5, 9, 14, 20, 27, 33, 41, 52, 60, 75)]
    print("case batch tile", tile, [p[0] for p in picks], flush=True)
    cs = [ctx.from_lists(offs, pos, u, bitmap_policy=(k % 3) | BUILD_ALL)
          for k, (name, u, offs, pos, _, _) in enumerate(picks)]
    try:
        with ctx.batch(cs) as b:
            for pol, graph in (("analytic", False), ("analytic", True), ("fixed", False),
                               ("analytic", True), ("learned", False)):
                st, sums, lat = b.allpairs(policy=pol, fixed_cell=4, out="tri", tile=tile,
                                           use_graph=graph)
                for c, (name, u, offs, pos, want_sum, want), s, l in zip(cs, picks, sums, lat):
                    assert int(s) == want_sum, (name, pol, graph)
                    assert np.array_equal(c.tri(), want), (name, pol, graph)
                    assert l >= 0.0
                assert st["sum"] == sum(p[4] for p in picks)
                assert st["pairs"] == sum(c.n_pairs for c in cs)
                assert sum(st["cell_pairs"].values()) == st["pairs"]
                ctx.sync()
            st, sums, _ = b.allpairs(policy="analytic", out="sum", tile=tile)
            assert [int(s) for s in sums] == [p[4] for p in picks]
    finally:
        for c in cs:
            c.free()


def test_batch_rejects_dense_members(sh, ctx):
    words = O.bernoulli_rows(1, 0.5, 128, 0, 4)
    with ctx.upload_contig(words) as d:
        with pytest.raises(sh.StormHipError):
            ctx.batch([d])


def test_autotune_general_keeps_counts(ctx, cases):
    name, u, offs, pos, want_sum, want = cases[20]
    print("case autotune", name, flush=True)
    with ctx.from_lists(offs, pos, u, bitmap_policy=1 | BUILD_ALL) as c:
        r = ctx.autotune(c, policy="analytic", tile=64)
        assert r["tile"] in (64, 128, 256)
        # 3 tile widths, then at 256 rows the launch shapes: (8 waves x stage
        # words {16, 8, 4} + 4 waves x {8, 4}) x band {0, 2, 4, 8} x K-split
        # target {1, 2, 4} (one target when the corpus has at least 4 tiles per CU)
        assert r["n_tried"] in (3, 3 + 20, 3 + 60)
        st = c.allpairs(policy="analytic", out="tri", tile=r["tile"], use_graph=True)
        assert st["sum"] == want_sum
        assert np.array_equal(c.tri(), want)
    ctx.set_bb_config(16, 4, 2, 8)


def test_bs_wide_stripes(ctx):
    """Wide B x S stripes: 1,300 rows (tiles 0..4 full, tile 5 of 20 rows) with
    cardinalities spanning 1..20,000 over a universe of 60,001, every row a
    bitmap, so the list tiles 0 and 1 run against the full bitmap tile pairs
    (2, 3) as wide stripes (one 64-byte slice load per element for 512 rows)
    and against (4, 5) as ordinary ones (tile 5 partial); also with bitmaps for
    the dense rows only (no wide stripe where a tile has lists). Exact per pair
    under the forced cell, the analytic policy, thresholded calls and an
    exploration call (its labels split a wide stripe's time over its two tile
    pairs)."""
    rng = np.random.default_rng(17)
    u = 60_001
    rows = [corpora.uniform_row(rng, u, int(np.exp(rng.random() * np.log(20_000))))
            for _ in range(1300)]
    rows[3] = np.zeros(0, np.uint32)
    offs, pos = corpora.csr(rows)
    want_sum, want = O.allpairs_lists(offs, pos)
    card = np.diff(offs.astype(np.int64)).astype(np.float64)
    I, J = np.triu_indices(len(rows), 1)
    v = want.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        jac = v / (card[I] + card[J] - v)
    for bitmaps in (1, 0):
        print("case bs wide stripes bitmaps", bitmaps, flush=True)
        with ctx.from_lists(offs, pos, u, bitmap_policy=bitmaps | BUILD_ALL | 64) as c:
            for pol in ("fixed", "analytic"):
                for graph in (False, True):
                    st = c.allpairs(policy=pol, fixed_cell=CELLS.index("BS"), out="tri",
                                    tile=256, use_graph=graph)
                    assert st["sum"] == want_sum, (bitmaps, pol, graph)
                    assert np.array_equal(c.tri(), want), (bitmaps, pol, graph)
                    assert sum(st["cell_pairs"].values()) == st["pairs"], (bitmaps, pol, st)
            st, hits = c.allpairs_threshold(0.05, policy="fixed", fixed_cell=CELLS.index("BS"),
                                            tile=256)
            ok = (want > 0) & (jac >= 0.05)
            exp = np.stack([I[ok], J[ok], want[ok]], 1).astype(np.uint32)
            assert np.array_equal(hits, exp), bitmaps
            X, Y, M, idx, s = ctx.selector_explore(c, tile=256, seed=3, fraction=1.0)
            assert s == want_sum, bitmaps
            T = (len(rows) + 255) // 256
            assert len(X) == T * (T + 1) // 2, (bitmaps, len(X))
            assert np.all(M.sum(axis=1) == 1), bitmaps


@pytest.mark.parametrize("tile", [64, 128, 256])
def test_bs_slices(ctx, tile):
    """B x S through bit-sliced bitmap groups: 300 rows whose cardinalities span
    1..40,000 over a universe of 100,003 (five 64-row groups, the last one
    partial), every row a bitmap or only the dense ones (then only the groups
    of dense rows have slices and the rest runs per pair); exact against the
    oracle under the forced cell and under the selecting policies."""
    rng = np.random.default_rng(91)
    u = 100_003
    rows = [corpora.uniform_row(rng, u, int(np.exp(rng.random() * np.log(40_000))))
            for _ in range(300)]
    rows[7] = np.zeros(0, np.uint32)
    rows[100] = np.arange(u, dtype=np.uint32)
    offs, pos = corpora.csr(rows)
    want_sum, want = O.allpairs_lists(offs, pos)
    for bitmaps in (1, 0):
        print("case bs slices tile", tile, "bitmaps", bitmaps, flush=True)
        with ctx.from_lists(offs, pos, u, bitmap_policy=bitmaps | BUILD_ALL | 64) as c:
            for pol in ("fixed", "analytic"):
                st = c.allpairs(policy=pol, fixed_cell=CELLS.index("BS"), out="tri", tile=tile)
                assert st["sum"] == want_sum, (bitmaps, pol)
                assert np.array_equal(c.tri(), want), (bitmaps, pol)
                assert sum(st["cell_pairs"].values()) == st["pairs"], (bitmaps, pol, st)
            # thresholded calls report through the same stripes
            st, hits = c.allpairs_threshold(0.2, policy="fixed", fixed_cell=CELLS.index("BS"),
                                            tile=tile)
            card = np.diff(offs.astype(np.int64)).astype(np.float64)
            I, J = np.triu_indices(len(rows), 1)
            v = want.astype(np.float64)
            with np.errstate(divide="ignore", invalid="ignore"):
                ok = (want > 0) & (v / (card[I] + card[J] - v) >= 0.2)
            exp = np.stack([I[ok], J[ok], want[ok]], 1).astype(np.uint32)
            assert np.array_equal(hits, exp), bitmaps


def _phase_rows(rng, u, shift, phases, card):
    """Positions whose bucket (p >> shift) falls in one of `phases` mod 6."""
    nb = (u >> shift) + 1
    buckets = np.array([b for b in range(nb) if b % 6 in phases], np.int64)
    pick = rng.choice(len(buckets), size=card, replace=True)
    p = buckets[pick] * (1 << shift) + rng.integers(0, 1 << shift, size=card)
    return np.unique(p[p < u]).astype(np.uint32)


def test_zone_maps(ctx):
    """Coarse zone maps (16,384 buckets; here bucket = position >> 6): twelve
    64-row groups with disjoint cardinality bands, so each 64-row tile of the
    density order is one group; groups 0..5 hold positions only in buckets of
    one residue mod 6, groups 6..11 in two residues. Tile pairs of disjoint
    residue sets are settled empty before any cell runs; tile pairs that
    share under half the buckets test every pair (rows of different residues
    are skipped). Exact per pair under the analytic policy and every per-pair
    cell, with more skipped pairs and empty tiles than without zone maps."""
    rng = np.random.default_rng(5)
    u = 1_000_003
    rows = []
    for g in range(12):
        ph = {g % 6} if g < 6 else {g % 6, (g + 1) % 6}
        for _ in range(64):
            rows.append(_phase_rows(rng, u, 6, ph, 40 + 100 * g + int(rng.integers(0, 50))))
    # bands are disjoint after de-duplication too: check the grouping holds
    cards = [len(r) for r in rows]
    assert all(max(cards[64 * g:64 * g + 64]) < min(cards[64 * g + 64:64 * g + 128])
               for g in range(11))
    offs, pos = corpora.csr(rows)
    want_sum, want = O.allpairs_lists(offs, pos)
    skipped = {}
    for zones in (0, 128):
        for bitmaps in (1, 0):
            print("case zone maps zones", zones, "bitmaps", bitmaps, flush=True)
            with ctx.from_lists(offs, pos, u, bitmap_policy=bitmaps | BUILD_ALL | zones) as c:
                for tile in (64, 256):
                    st = c.allpairs(policy="analytic", out="tri", tile=tile)
                    assert st["sum"] == want_sum, (zones, bitmaps, tile)
                    assert np.array_equal(c.tri(), want), (zones, bitmaps, tile)
                    if tile == 64 and bitmaps == 1:
                        empties = st["cell_tiles"].get("EMPTY", 0)
                        if zones:
                            assert empties > 0, st["cell_tiles"]
                        else:
                            assert empties == 0, st["cell_tiles"]
                cells = ["SS", "SR", "SW", "RR", "RW", "WW"] + (["BS", "BR", "BW"] if bitmaps else [])
                for cell in cells:
                    st = c.allpairs(policy="fixed", fixed_cell=CELLS.index(cell), out="tri",
                                    tile=64)
                    assert st["sum"] == want_sum, (zones, bitmaps, cell)
                    assert np.array_equal(c.tri(), want), (zones, bitmaps, cell)
                    assert sum(st["cell_pairs"].values()) == st["pairs"], (cell, st)
                    skipped[(zones, bitmaps, cell)] = st["skipped"]
    for (zones, bitmaps, cell), sk in skipped.items():
        if zones:
            assert sk > skipped[(0, bitmaps, cell)], (bitmaps, cell, sk)


def test_bb_zone_map_plan_per_pair(ctx):
    """bb_occ_sel for one pair: 300 dense rows (1,000-1,400 positions, so
    every row carries a zone map of 8-word bins) clustered in three windows of
    4,096 positions each over a universe of 300,007; under PerPair most pairs
    choose B x B, priced by their live bins, and count only the bins live on
    both sides (or none). Exact against the oracle."""
    rng = np.random.default_rng(23)
    u = 300_007
    rows = []
    for r in range(300):
        starts = rng.integers(0, u - 4096, size=3)
        k = int(rng.integers(1000, 1400))
        p = np.concatenate([s + rng.integers(0, 4096, size=k // 3 + 1) for s in starts])
        rows.append(np.unique(p).astype(np.uint32))
    offs, pos = corpora.csr(rows)
    want_sum, want = O.allpairs_lists(offs, pos)
    with ctx.from_lists(offs, pos, u, bitmap_policy=1 | BUILD_ALL) as c:
        for tile in (64, 256):
            st = c.allpairs(policy="per_pair", out="tri", tile=tile)
            assert st["sum"] == want_sum, tile
            assert np.array_equal(c.tri(), want), tile
            assert st["cell_pairs"].get("BB", 0) > 0, st["cell_pairs"]


def _row_sums(n, tri):
    I, J = np.triu_indices(n, 1)
    rs = np.zeros(n, np.uint64)
    np.add.at(rs, I, tri.astype(np.uint64))
    np.add.at(rs, J, tri.astype(np.uint64))
    return rs


def test_density_order_triangle(ctx, cases):
    """out="tri_density": the triangle packed over positions of the density
    order (the coordinates of the reference's allpairs_tiles under PerTile).
    Rearranged by the corpus's order it equals the original-order triangle of
    the oracle, for the selecting policies (B x S stripes, wide ones
    included, joins and B x B tiles writing by position; probe calls with
    their shadow instances, refined and per-pair tiles; the learned policy,
    as analytic where no selector is trained), forced cells, the all-bitmap
    path and batches; row sums come back by original row."""
    for name, u, offs, pos, want_sum, want in cases[::6]:
        print("case density tri", name, flush=True)
        n = len(offs) - 1
        rs_want = _row_sums(n, want) if n > 1 else None
        with ctx.from_lists(offs, pos, u, bitmap_policy=1 | BUILD_ALL | 64) as c:
            runs = [("analytic", 0, 64, False), ("analytic", 0, 256, True),
                    ("fixed", CELLS.index("BS"), 256, False), ("fixed", CELLS.index("SS"), 64, False),
                    ("fixed", CELLS.index("RR"), 64, True), ("fixed", CELLS.index("BB"), 256, False),
                    ("all_bitmap", 0, 64, False), ("probe", 0, 64, False), ("refine", 0, 64, False),
                    ("per_pair", 0, 64, False), ("learned", 0, 256, False)]
            for pol, cell, tile, graph in runs:
                st = c.allpairs(policy=pol, fixed_cell=cell, out="tri_density", tile=tile,
                                use_graph=graph)
                assert st["sum"] == want_sum, (name, pol, cell, tile)
                got = c.tri()
                assert np.array_equal(c.tri_to_original(got), want), (name, pol, cell, tile)
                if n > 1:
                    assert np.array_equal(c.row_sums(), rs_want), (name, pol, cell, tile)
            # The original-order call after a density-order one (the cached
            # plan is rebuilt for the other packing).
            st = c.allpairs(policy="analytic", out="tri", tile=256, use_graph=True)
            assert np.array_equal(c.tri(), want), name


def test_density_order_triangle_wide_stripes_and_batch(ctx, cases):
    """The density-packed triangle through wide B x S stripes (1,300 rows of
    cardinality 1..20,000, every row a bitmap) and through a batch."""
    rng = np.random.default_rng(17)
    u = 60_001
    rows = [corpora.uniform_row(rng, u, int(np.exp(rng.random() * np.log(20_000))))
            for _ in range(1300)]
    offs, pos = corpora.csr(rows)
    want_sum, want = O.allpairs_lists(offs, pos)
    print("case density tri wide stripes", flush=True)
    with ctx.from_lists(offs, pos, u, bitmap_policy=1 | BUILD_ALL | 64) as c:
        for pol in ("fixed", "analytic"):
            st = c.allpairs(policy=pol, fixed_cell=CELLS.index("BS"), out="tri_density",
                            tile=256, use_graph=True)
            assert st["sum"] == want_sum, pol
            assert np.array_equal(c.tri_to_original(c.tri()), want), pol
        assert np.array_equal(c.row_sums(), _row_sums(len(rows), want))
    picks = [cases[k] for k in (0, 5, 14, 27, 41, 60)]
    print("case density tri batch", [p[0] for p in picks], flush=True)
    cs = [ctx.from_lists(offs_, pos_, u_, bitmap_policy=(k % 3) | BUILD_ALL | 64)
          for k, (name, u_, offs_, pos_, _, _) in enumerate(picks)]
    try:
        with ctx.batch(cs) as b:
            for tile, graph in ((64, False), (256, True)):
                st, sums, _ = b.allpairs(policy="analytic", out="tri_density", tile=tile,
                                         use_graph=graph)
                for c, (name, u_, offs_, pos_, want_sum_, want_) in zip(cs, picks):
                    assert np.array_equal(c.tri_to_original(c.tri()), want_), (name, tile)
                assert [int(s) for s in sums] == [p[4] for p in picks]
    finally:
        for c in cs:
            c.free()
