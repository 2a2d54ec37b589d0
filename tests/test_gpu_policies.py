"""The measuring and per-pair policies on the GPU (Policy::Probe, Refine,
PerPair / Oracle; kernels/storm_allpairs.cpp:445-482, 563-588, 636-660):
exact counts under every tile width and output form, and the per-pair cell
choices against the oracle's select_pairing / predict on the rows' own
metadata."""
import ctypes as C
import zlib

import numpy as np
import pytest

import corpora
import oracle_py as O

pytestmark = pytest.mark.gpu

BUILD_ALL = 4 | 8 | 16
EMPTY = 10
POLICIES = ["probe", "refine", "per_pair"]


@pytest.fixture(scope="module")
def fresh(sh):
    """A context of its own: the default cost model, whatever ran before."""
    c = sh.Context(0)
    yield c
    c.close()


def _case(name, n, u, d, st, sp):
    rows = corpora.mixed_corpus(zlib.crc32(name.encode()), n, u, d, st, sp)
    offs, pos = corpora.csr(rows)
    s, tri = O.allpairs_lists(offs, pos)
    return name, u, offs, pos, s, tri


def _hetero(name, n_tiny, n_dense, u, lo, hi, runs):
    """Tiny rows (1-3 elements) then a few larger ones, so that a
    density-sorted tile holds mostly tiny rows and a handful of large ones:
    the heterogeneous tiles the refine gate is about."""
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    rows = []
    for k in range(n_tiny + n_dense):
        c = int(rng.integers(1, 4)) if k < n_tiny else int(rng.integers(lo, hi))
        rows.append(corpora.runs_row(rng, u, c, 24) if runs else
                    np.sort(rng.choice(u, size=c, replace=False)).astype(np.uint32))
    offs, pos = corpora.csr(rows)
    s, tri = O.allpairs_lists(offs, pos)
    return name, u, offs, pos, s, tri


@pytest.fixture(scope="module")
def cases():
    out = [_hetero("hetero-uniform", 189, 67, 16384, 150, 400, False),
           _hetero("hetero-runs", 189, 67, 65536, 200, 2000, True),
           _case("bimodal", 640, 16384, 0.02, "uniform", "bimodal"),
           _case("runs-inverse", 520, 65536, 0.01, "runs", "inverse"),
           _case("dense", 300, 4033, 0.3, "uniform", "inverse")]
    for name, rows, u in corpora.all_corpora()[::9]:
        offs, pos = corpora.csr(rows)
        s, tri = O.allpairs_lists(offs, pos)
        out.append((name, u, offs, pos, s, tri))
    return out


@pytest.mark.parametrize("policy", POLICIES)
def test_exact_every_tile_width(fresh, cases, policy):
    for name, u, offs, pos, want_sum, want in cases:
        for bitmaps in (0, 1):
            with fresh.from_lists(offs, pos, u, bitmap_policy=bitmaps | BUILD_ALL) as c:
                for tile in (64, 128, 256):
                    print("case", policy, name, bitmaps, tile, flush=True)
                    st = c.allpairs(policy=policy, out="tri", tile=tile)
                    assert st["sum"] == want_sum, (name, tile)
                    assert np.array_equal(c.tri(), want), (name, tile)
                    assert sum(st["cell_pairs"].values()) == st["pairs"], (name, tile, st)


def test_probe_measures_then_commits(fresh, cases):
    """Sampled tiles run both candidates (the runner-up without output), the
    rest are committed per class; sums stay exact, eager and graph alike."""
    name, u, offs, pos, want_sum, want = cases[2]
    with fresh.from_lists(offs, pos, u, bitmap_policy=BUILD_ALL) as c:
        for tile in (64, 256):
            st = c.allpairs(policy="probe", out="tri", tile=tile)
            print("probe", tile, st["probe_tiles"], st["probe_switched"], st["cell_tiles"],
                  flush=True)
            assert st["probe_tiles"] > 0, st
            assert st["sum"] == want_sum and np.array_equal(c.tri(), want)
            for rep in range(3):
                sg = c.allpairs(policy="probe", out="tri", tile=tile, use_graph=True)
                assert sg["sum"] == want_sum, rep
                assert np.array_equal(c.tri(), want), rep
                assert sg["probe_tiles"] > 0
            # The shadow runs leave the pair counts alone.
            assert sum(sg["cell_pairs"].values()) == sg["pairs"]


def _row_metas(rows, u):
    class M(C.Structure):
        _fields_ = [("a", C.c_uint32 * 6), ("hr", C.c_uint8), ("ho", C.c_uint8)]

    metas = (M * len(rows))()
    for i, r in enumerate(rows):
        k = len(r)
        runs = int(np.count_nonzero(np.diff(r.astype(np.int64)) != 1) + 1) if k else 0
        metas[i].a[0], metas[i].a[1], metas[i].a[2] = k, runs, (u + 63) // 64
        metas[i].a[3] = len(np.unique(r >> 6))
        metas[i].a[4] = int(r[0]) if k else 0xFFFFFFFF
        metas[i].a[5] = int(r[-1]) if k else 0
        metas[i].hr = int(runs and k // runs >= 512)
        metas[i].ho = int(k * 512 > u)
    return M, metas


def _disjoint(a, b):
    return not len(a) or not len(b) or a[-1] < b[0] or b[-1] < a[0]


def test_per_pair_cells_are_select_pairing_per_pair(fresh, cases):
    """PER_PAIR: every pair not settled by the span test runs
    select_pairing(denser row, other row) (storm_cost.cpp:422-465 on the rows'
    RowMeta): the device's per-cell pair counts equal the oracle's."""
    L = O.oracle()
    model = (C.c_double * 16)()
    L.orc_default_model(model)
    for name, u, offs, pos, want_sum, want in cases[:5]:
        rows = [pos[int(offs[i]):int(offs[i + 1])] for i in range(len(offs) - 1)]
        M, metas = _row_metas(rows, u)
        hist = np.zeros(12, np.int64)
        for i in range(len(rows)):
            for j in range(i + 1, len(rows)):
                if _disjoint(rows[i], rows[j]):
                    hist[EMPTY] += 1
                    continue
                d, s = (i, j) if len(rows[i]) >= len(rows[j]) else (j, i)
                hist[L.orc_select_pairing(model, C.byref(metas[d]), C.byref(metas[s]), None)] += 1
        with fresh.from_lists(offs, pos, u, bitmap_policy=1 | BUILD_ALL) as c:
            st = c.allpairs(policy="per_pair", out="tri", tile=64)
        assert st["sum"] == want_sum
        got = np.zeros(12, np.int64)
        for k, v in st["cell_pairs"].items():
            got[corpora_cells().index(k)] = v
        assert np.array_equal(got, hist), (name, got, hist)


def corpora_cells():
    return ["BB", "BS", "BR", "BW", "SS", "SR", "SW", "RR", "RW", "WW", "EMPTY", "CB"]


def test_refine_cells_follow_the_gate(fresh, cases):
    """REFINE: on a tile whose rows spread by max/mean cardinality >= 8 and
    whose runner-up is predicted within 3x, each pair takes the cheaper of
    the two as priced on its own rows; other tiles run the best candidate
    (B x B tiles count every pair, the others settle disjoint pairs as EMPTY).
    The device's per-cell pair counts equal this reconstruction."""
    L = O.oracle()
    model = (C.c_double * 16)()
    L.orc_default_model(model)

    def pred(p, a, b):
        return L.orc_predict(model, p, C.byref(a), C.byref(b), 0.05, 16.0, 1.0)

    refined_any = False
    for name, u, offs, pos, want_sum, want in cases[:5]:
        rows = [pos[int(offs[i]):int(offs[i + 1])] for i in range(len(offs) - 1)]
        n = len(rows)
        M, metas = _row_metas(rows, u)
        order = np.zeros(n, np.uint32)
        L.orc_density_order(metas, n, order.ctypes.data_as(C.c_void_p))
        tile = 64
        T = (n + tile - 1) // tile
        hist = np.zeros(12, np.int64)
        for bi in range(T):
            for bj in range(bi, T):
                ia = order[bi * tile:min(n, bi * tile + tile)]
                ib = order[bj * tile:min(n, bj * tile + tile)]
                ta, tb = M(), M()
                L.orc_tile_meta(metas, order.ctypes.data_as(C.c_void_p), bi * tile,
                                min(n, bi * tile + tile), C.byref(ta))
                L.orc_tile_meta(metas, order.ctypes.data_as(C.c_void_p), bj * tile,
                                min(n, bj * tile + tile), C.byref(tb))
                pairs = [(int(x), int(y)) for xi, x in enumerate(ia) for yi, y in enumerate(ib)
                         if bi != bj or yi > xi]
                ka = [rows[x] for x in ia if len(rows[x])]
                kb = [rows[y] for y in ib if len(rows[y])]
                if (not ka or not kb or max(int(r[-1]) for r in ka) < min(int(r[0]) for r in kb)
                        or max(int(r[-1]) for r in kb) < min(int(r[0]) for r in ka)):
                    hist[EMPTY] += len(pairs)
                    continue
                first = L.orc_select_pairing(model, C.byref(ta), C.byref(tb), None)
                # tile_top2 gives no runner-up on the emptiness proof (B x B).
                proof = first == EMPTY
                if proof:
                    first = 0
                c1 = pred(first, ta, tb)
                second, c2 = first, 0.0
                for k in range(10):
                    if proof or k == first:
                        continue
                    ck = pred(k, ta, tb)
                    if second == first or ck < c2:
                        second, c2 = k, ck
                ratio = c2 / c1 if c1 > 0 and second != first else 1e300
                cards_a = [len(rows[x]) for x in ia]
                cards_b = [len(rows[y]) for y in ib]
                mean_a, mean_b = sum(cards_a) / len(ia), sum(cards_b) / len(ib)
                het = max(max(cards_a) / mean_a if mean_a > 0 else 1.0,
                          max(cards_b) / mean_b if mean_b > 0 else 1.0)
                gate = second != first and het >= 8.0 and ratio <= 3.0
                for x, y in pairs:
                    if first == 0 and not gate:
                        hist[0] += 1
                        continue
                    if _disjoint(rows[x], rows[y]):
                        hist[EMPTY] += 1
                        continue
                    if not gate:
                        hist[first] += 1
                        continue
                    refined_any = True
                    d, s = (x, y) if len(rows[x]) >= len(rows[y]) else (y, x)
                    hist[second if pred(second, metas[d], metas[s]) <
                         pred(first, metas[d], metas[s]) else first] += 1
        with fresh.from_lists(offs, pos, u, bitmap_policy=1 | BUILD_ALL) as c:
            st = c.allpairs(policy="refine", out="tri", tile=tile)
        assert st["sum"] == want_sum
        got = np.zeros(12, np.int64)
        for k, v in st["cell_pairs"].items():
            got[corpora_cells().index(k)] = v
        assert np.array_equal(got, hist), (name, got, hist)
    assert refined_any, "no tile passed the refine gate: the corpora do not exercise it"


@pytest.mark.parametrize("policy", POLICIES)
def test_threshold_band_and_batch(fresh, cases, policy):
    """The same policies in thresholded, band and batched calls."""
    name, u, offs, pos, want_sum, want = cases[2]
    n = len(offs) - 1
    with fresh.from_lists(offs, pos, u, bitmap_policy=BUILD_ALL) as c:
        _, hits_a = c.allpairs_threshold(0.05, policy="analytic", tile=64)
        for tile in (64, 256):
            _, hits = c.allpairs_threshold(0.05, policy=policy, tile=tile)
            assert np.array_equal(hits, hits_a), (policy, tile)
        m = np.zeros((n, n), np.uint32)
        m[np.triu_indices(n, 1)] = want
        m = m + m.T
        perm = c.order()
        for band in (0, 4, (n - 1) // 128):
            _, got = c.allpairs_band(band, policy=policy, tile=128)
            i0, i1 = band * 128, min(n, band * 128 + 128)
            exp = m[np.ix_(perm[i0:i1], perm)]
            exp = np.where(np.arange(n)[None, :] > np.arange(i0, i1)[:, None], exp, 0)
            assert np.array_equal(got, exp), (policy, band)
    members = [fresh.from_lists(o, p, uu, bitmap_policy=BUILD_ALL)
               for _, uu, o, p, _, _ in cases[:3]]
    try:
        b = fresh.batch(members)
        st, sums, _ = b.allpairs(policy=policy, tile=64)
        assert list(sums) == [s for _, _, _, _, s, _ in cases[:3]], policy
        b.free()
    finally:
        for c in members:
            c.free()
