"""Pins the CPU oracle (oracle/storm_oracle.c) before anything is checked
against it: against the compiled reference (oracle/_ref) where it is built,
against closed forms, brute force and the reference's own known answers."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O

P = O._p
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference checker not built")


def lists_csr(rows):
    offs = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.uint64)
    pos = np.concatenate([np.asarray(r, dtype=np.uint32) for r in rows]) if rows else \
        np.zeros(0, np.uint32)
    return offs, pos.astype(np.uint32)


# ---------------------------------------------------------------- stream
@needs_ref
def test_rng_matches_reference_stream():
    L, R = O.oracle(), O.reference()
    for seed in (0, 1, 2, 5, 0xDEADBEEF):
        want = np.zeros(64, dtype=np.uint64)
        R.ref_rng_next(seed, 64, P(want))
        s = C.c_uint64(seed if seed else 0x9E3779B97F4A7C15)
        got = [L.orc_rng_next(C.byref(s)) for _ in range(64)]
        assert list(map(int, want)) == got


def test_bernoulli_row_skip_is_the_same_stream():
    full = O.bernoulli_rows(9, 0.3, 1000, 0, 7)
    part = O.bernoulli_rows(9, 0.3, 1000, 3, 4)
    assert np.array_equal(full[3:], part)
    # bits past n_bits stay zero
    assert not np.any(full[:, -1] >> np.uint64(1000 - 960))


def test_config1_known_answer():
    """Config 1 (N=2048 x M=262,144, Bernoulli(0.01), Rng(1)): the reference
    CPU paths all give 54,935,083; the closed form sum_b C(c_b,2) must too."""
    words = O.bernoulli_rows(1, 0.01, 262144, 0, 2048)
    total, rs = O.column_closed_forms(words)
    assert total == 54_935_083
    assert int(rs.sum()) == 2 * total


# ---------------------------------------------------------------- per-pair
def test_intersect_three_ways_agree():
    rng = np.random.default_rng(1)
    for u in (64, 1000, 4033, 70001):
        for _ in range(20):
            a = np.unique(rng.integers(0, u, rng.integers(0, 200))).astype(np.uint32)
            b = np.unique(rng.integers(0, u, rng.integers(0, 200))).astype(np.uint32)
            nw = (u + 63) // 64
            L = O.oracle()
            x = L.orc_intersect_lists(P(a), len(a), nw, P(b), len(b), nw)
            y = L.orc_intersect_merge(P(a), len(a), P(b), len(b))
            assert x == y == len(np.intersect1d(a, b))


def test_bitloop_popcount_matches_builtin():
    L = O.oracle()
    rng = np.random.default_rng(2)
    for v in [0, 1, 2**63, 2**64 - 1] + [int(x) for x in rng.integers(0, 2**63, 50)]:
        assert L.orc_popcount_bitloop(v) == bin(v).count("1")


def test_allpairs_words_and_lists_agree_with_brute_force():
    rng = np.random.default_rng(3)
    rows = [np.unique(rng.integers(0, 300, rng.integers(0, 60))).astype(np.uint32)
            for _ in range(23)]
    offs, pos = lists_csr(rows)
    s1, t1 = O.allpairs_lists(offs, pos)
    words = np.zeros((23, 5), dtype=np.uint64)
    for i, r in enumerate(rows):
        for p in r:
            words[i, p >> 6] |= np.uint64(1) << np.uint64(p & 63)
    s2, t2 = O.allpairs_words(words)
    brute = [len(np.intersect1d(rows[i], rows[j])) for i in range(23) for j in range(i + 1, 23)]
    assert s1 == s2 == sum(brute)
    assert list(t1) == brute and list(t2) == brute
    total, rs = O.column_closed_forms(words)
    assert total == s1
    rs_tri = np.zeros(23, dtype=np.uint64)
    O.oracle().orc_row_sums_from_tri(P(t1), 23, P(rs_tri))
    assert np.array_equal(rs, rs_tri)
    t3, rs3 = O.lists_closed_forms(offs, pos, 300)
    assert t3 == total and np.array_equal(rs3, rs)


# ---------------------------------------------------------------- layouts
EDGE_UNIVERSES = [64, 128, 1024, 1025, 4033, 5000, 8192, 65536, 70001]


def edge_rows(u):
    """tests/test_cells.cpp:195-230 shapes."""
    rows = [[], [0], [u - 1], [63, 64], [511, 512], list(range(u)), list(range(0, u, 64)),
            list(range(0, u, 2)), list(range(u // 2, min(u // 2 + 300, u))),
            list(range(1, min(200, u)))]
    return [np.array(sorted({p for p in r if p < u}), dtype=np.uint32) for r in rows]


ref_row = O.ref_build_row


orc_row = O.build_row


def _rowsets():
    out = []
    for u in EDGE_UNIVERSES:
        for r in edge_rows(u):
            out.append((u, r))
    rng = np.random.default_rng(4)
    for u in (1024, 4033, 16384, 65536):
        for d in (5e-4, 0.02, 0.3, 0.7):
            k = max(0, int(d * u))
            out.append((u, np.unique(rng.integers(0, u, k)).astype(np.uint32)))
        # long runs so the rank gate (mean run >= 512) opens
        out.append((u, np.arange(u // 4, u // 4 + min(u // 2, 3000), dtype=np.uint32)))
    return out


@needs_ref
def test_build_row_matches_reference_layouts():
    for u, pos in _rowsets():
        want = ref_row(pos, u)
        got = orc_row(pos, u)
        assert np.array_equal(got["meta"], want["meta"]), (u, len(pos))
        assert got["has_rank"] == want["sizes"][7] and got["has_occ"] == want["sizes"][8]
        for k in ("bitmap", "occ", "run_start", "run_end", "ewah", "comp"):
            assert np.array_equal(got[k], want[k]), (k, u, len(pos))
        if want["sizes"][1]:
            # the reference stores 2*(nblk+1) words; the sentinel's pair word is 0
            assert np.array_equal(got["rank"], want["rank"]), (u, len(pos))


def test_representation_round_trips():
    """tests/test_cells.cpp:81-152: S->B, R->B (maximal), W->B, rank at every
    word and interior bits, complement."""
    L = O.oracle()
    for u, pos in _rowsets():
        r = orc_row(pos, u)
        nw = (u + 63) // 64
        b = np.zeros(nw, np.uint64)
        for p in pos:
            b[p >> 6] |= np.uint64(1) << np.uint64(p & 63)
        assert np.array_equal(b, r["bitmap"])
        rb = np.zeros(u + 1, bool)
        for s, e in zip(r["run_start"], r["run_end"]):
            rb[s:e] = True
        assert np.array_equal(np.flatnonzero(rb), pos.astype(np.int64))
        assert np.all(r["run_start"][1:] > r["run_end"][:-1])
        dec = np.zeros(nw + 1, np.uint64)
        assert L.orc_ewah_decode(P(r["ewah"]), len(r["ewah"]), P(dec)) == nw
        assert np.array_equal(dec[:nw], b)
        nblk = (nw + 7) // 8
        rank = np.zeros(2 * (nblk + 1), np.uint64)
        L.orc_build_rank(P(b), nw, P(rank))
        L.orc_rank_at.restype = C.c_uint64
        L.orc_rank_at.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32]
        cum = np.concatenate([[0], np.cumsum(rb[:u])])
        for x in list(range(0, u, 64)) + [min(u - 1, 5), min(u - 1, 77), u - 1, u]:
            assert L.orc_rank_at(P(b), nw, P(rank), x) == cum[min(x, u)]
        if 2 * len(pos) > u:
            assert np.array_equal(r["comp"], np.setdiff1d(np.arange(u), pos))


# ---------------------------------------------------------------- cost model
def _metas_from(rows, u):
    out = []
    for r in rows:
        o = orc_row(r, u)
        out.append(np.concatenate([o["meta"], [o["has_rank"], o["has_occ"]]]).astype(np.uint32))
    return out


def _random_metas(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nw = int(rng.choice([16, 1000, 60420, 577728]))
        card = int(rng.integers(0, nw * 64))
        runs = int(rng.integers(0 if card == 0 else 1, max(2, card + 1)))
        nnz = int(min(nw, rng.integers(0, card + 1)))
        first = int(rng.integers(0, nw * 64))
        last = int(rng.integers(first, nw * 64))
        out.append(np.array([card, runs, nw, nnz, first, last, rng.integers(0, 2),
                             rng.integers(0, 2)], np.uint32))
    return out


@needs_ref
def test_select_pairing_and_predict_match_reference():
    L, R = O.oracle(), O.reference()

    class M(C.Structure):
        _fields_ = [("a", C.c_uint32 * 6), ("hr", C.c_uint8), ("ho", C.c_uint8)]

    def as_orc(m):
        x = M()
        for k in range(6):
            x.a[k] = int(m[k])
        x.hr, x.ho = int(m[6]), int(m[7])
        return x

    model = (C.c_double * 16)()
    L.orc_default_model(model)
    metas = _random_metas(300, 5) + _metas_from(
        [r for u in (1024, 65536) for r in edge_rows(u)], 65536)
    rng = np.random.default_rng(6)
    for _ in range(3000):
        a, b = metas[rng.integers(len(metas))], metas[rng.integers(len(metas))]
        pa, pb = C.c_double(0), C.c_double(0)
        ga, gb = as_orc(a), as_orc(b)
        got = L.orc_select_pairing(model, C.byref(ga), C.byref(gb), C.byref(pa))
        want = R.ref_select_pairing(P(a), P(b), C.byref(pb))
        assert got == want, (a, b)
        assert pa.value == pytest.approx(pb.value, rel=1e-12, abs=1e-12)
        for p in range(11):
            assert L.orc_predict(model, p, C.byref(ga), C.byref(gb), 0.05, 16.0, 1.0) == \
                pytest.approx(R.ref_predict(p, P(a), P(b)), rel=1e-12, abs=1e-12)


def test_select_pairing_emptiness_proof():
    """storm_cost.cpp:427-431: empty rows or disjoint spans -> Empty (10)."""
    L = O.oracle()
    model = (C.c_double * 16)()
    L.orc_default_model(model)

    class M(C.Structure):
        _fields_ = [("a", C.c_uint32 * 6), ("hr", C.c_uint8), ("ho", C.c_uint8)]

    def mk(card, first, last):
        x = M()
        x.a[0], x.a[1], x.a[2], x.a[3], x.a[4], x.a[5] = card, 1, 100, 1, first, last
        return x

    assert L.orc_select_pairing(model, C.byref(mk(0, 0, 0)), C.byref(mk(5, 1, 9)), None) == 10
    assert L.orc_select_pairing(model, C.byref(mk(3, 0, 9)), C.byref(mk(3, 10, 20)), None) == 10
    assert L.orc_select_pairing(model, C.byref(mk(3, 0, 10)), C.byref(mk(3, 10, 20)), None) != 10


# ---------------------------------------------------------------- end to end
@needs_ref
@pytest.mark.parametrize("structure", [0, 1, 2])
@pytest.mark.parametrize("spectrum", [0, 1])
@pytest.mark.parametrize("u", [1024, 4033, 16384])
def test_allpairs_sum_matches_reference_policies(structure, spectrum, u):
    """tests/test_cells.cpp:237-287 sweep: reference allpairs_sum under every
    shipped policy equals the oracle sum over the reference's own corpus."""
    offs, pos = O.ref_generate(40, u, 0.02, structure, spectrum, 0x9911 + u)
    want, tri = O.allpairs_lists(offs, pos)
    R = O.reference()
    o32 = offs.astype(np.uint32)
    for policy in (0, 1, 2, 5):   # AllBitmap, PerPair, PerTile, Probe
        for tile in (8, 16, 64):
            cells = np.zeros(11, np.uint64)
            pairs = C.c_uint64(0)
            s = R.ref_allpairs_sum(P(o32), P(pos), 40, u, policy, tile, P(cells),
                                   C.byref(pairs), None)
            assert pairs.value == 780
            # The reference returns 0 for pairs routed to S x W / R x W
            # (storm_allpairs.cpp:245-249); elsewhere it is exact.
            if cells[6] == 0 and cells[8] == 0:
                assert s == want, (policy, tile)
    out = np.zeros(780, np.uint32)
    assert R.ref_allpairs_tiles_tri(P(o32), P(pos), 40, u, 64, P(out)) == want
    assert np.array_equal(out, tri)


# ---------------------------------------------------------------- MLP math
def _mlp_params(f, h1, h2, o, seed):
    rng = np.random.default_rng(seed)
    return [rng.normal(0, 0.5, s) for s in [(h1, f), (h1,), (h2, h1), (h2,), (o, h2), (o,)]]


def test_mlp_gradients_match_finite_differences():
    L = O.oracle()
    f, h1, h2, o, B = 5, 7, 6, 3, 9
    ps = _mlp_params(f, h1, h2, o, 7)
    rng = np.random.default_rng(8)
    x = rng.normal(size=(B, f))
    t = rng.normal(size=(B, o))
    mask = (rng.random((B, o)) < 0.8).astype(np.uint8)
    grads = [np.zeros_like(p) for p in ps]

    def loss(params, g=None):
        a = [P(np.ascontiguousarray(p)) for p in params]
        gg = [P(q) for q in g] if g else [None] * 6
        return L.orc_mlp_forward_backward(f, h1, h2, o, B, P(x), P(t), P(mask), *a, None, *gg)

    loss(ps, grads)
    eps = 1e-6
    for k, p in enumerate(ps):
        flat = p.reshape(-1)
        for idx in range(0, flat.size, max(1, flat.size // 7)):
            old = flat[idx]
            flat[idx] = old + eps
            lp = loss(ps)
            flat[idx] = old - eps
            lm = loss(ps)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            assert grads[k].reshape(-1)[idx] == pytest.approx(fd, rel=1e-5, abs=1e-8)


def test_adam_matches_published_update():
    """Kingma & Ba (2015) Algorithm 1 on a closed-form case: with a constant
    gradient g, m_t/(1-b1^t) = g and v_t/(1-b2^t) = g^2, so every step moves
    each parameter by -lr * g / (|g| + eps)."""
    L = O.oracle()
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -0.1, 2.0])
    m = np.zeros(3)
    v = np.zeros(3)
    for step in range(1, 6):
        L.orc_adam_step(3, P(p), P(g), P(m), P(v), step, 0.01, 0.9, 0.999, 1e-8)
    want = np.array([1.0, -2.0, 0.5]) - 5 * 0.01 * g / (np.abs(g) + 1e-8)
    assert np.allclose(p, want, rtol=0, atol=1e-12)


# ---------------------------------------------------------------- driver helpers
class OrcMeta(C.Structure):
    """orc_meta (storm_oracle.h): the six RowMeta integers, has_rank, has_occ."""
    _fields_ = [("a", C.c_uint32 * 6), ("hr", C.c_uint8), ("ho", C.c_uint8)]


def _orc_metas(offs, pos, u):
    n = len(offs) - 1
    arr = (OrcMeta * n)()
    for i in range(n):
        r = O.build_row(pos[int(offs[i]):int(offs[i + 1])], u)
        for k in range(6):
            arr[i].a[k] = int(r["meta"][k])
        arr[i].hr, arr[i].ho = int(r["has_rank"]), int(r["has_occ"])
    return arr


@needs_ref
def test_rng_below_matches_reference():
    """Rng::below (kernels/storm_gen.h:93-99), including bound 0 and bounds
    near 2^32 where the rejection limit is furthest from 2^64."""
    L, R = O.oracle(), O.reference()
    for seed in (0, 3, 0x9911, 0xFFFFFFFFFFFF):
        for bound in (0, 1, 2, 7, 1000, 3866847, (1 << 31) + 1, 3_000_000_000, (1 << 32) - 1):
            want = np.zeros(500, np.uint32)
            R.ref_rng_below(seed, bound, 500, P(want))
            s = C.c_uint64(L.orc_rng_seed(seed))
            got = [L.orc_rng_below(C.byref(s), bound) for _ in range(500)]
            assert np.array_equal(np.array(got, np.uint32), want), (seed, bound)
            if bound:
                assert int(want.max()) < bound


def test_density_order_is_a_stable_sort():
    """density_order (kernels/storm_allpairs.cpp:316-323) is std::stable_sort by
    cardinality: equal to Python's (stable) sorted() on every tie pattern."""
    L = O.oracle()
    rng = np.random.default_rng(17)
    for n in (0, 1, 2, 3, 17, 64, 65, 1000):
        for levels in (1, 3, 50, 10 ** 6):
            cards = rng.integers(0, levels, size=n)
            arr = (OrcMeta * max(n, 1))()
            for i in range(n):
                arr[i].a[0] = int(cards[i])
            order = np.zeros(max(n, 1), np.uint32)
            L.orc_density_order(arr, n, P(order))
            want = sorted(range(n), key=lambda i: int(cards[i]))
            assert list(order[:n]) == want, (n, levels)


def test_tile_meta_brute_force():
    """tile_meta (kernels/storm_allpairs.cpp:281-312) on random metadata:
    integer-truncated means, min first / max last over non-empty rows only
    (first 0 when all are empty), AND of the index flags, max n_words; an
    empty range keeps the flags false and divides by 1."""
    L = O.oracle()
    rng = np.random.default_rng(23)
    n = 300
    arr = (OrcMeta * n)()
    rows = []
    for i in range(n):
        card = int(rng.integers(0, 4)) * int(rng.integers(0, 1 << 20))
        first = int(rng.integers(0, 1 << 30))
        m = [card, int(rng.integers(0, 1 << 20)), int(rng.integers(1, 1 << 16)),
             int(rng.integers(0, 1 << 16)), first, first + int(rng.integers(0, 1 << 30))]
        hr, ho = int(rng.random() < 0.8), int(rng.random() < 0.8)
        for k in range(6):
            arr[i].a[k] = m[k]
        arr[i].hr, arr[i].ho = hr, ho
        rows.append((m, hr, ho))
    order = rng.permutation(n).astype(np.uint32)
    for lo, hi in [(0, 0), (5, 5), (0, 1), (0, 64), (10, 77), (250, 300), (0, 300)]:
        out = OrcMeta()
        L.orc_tile_meta(arr, P(order), lo, hi, C.byref(out))
        sel = [rows[int(order[k])] for k in range(lo, hi)]
        cnt = max(1, hi - lo)
        live = [m for m, _, _ in sel if m[0]]
        want = [sum(m[0] for m, _, _ in sel) // cnt, sum(m[1] for m, _, _ in sel) // cnt,
                max([m[2] for m, _, _ in sel], default=0), sum(m[3] for m, _, _ in sel) // cnt,
                min([m[4] for m in live], default=0), max([m[5] for m in live], default=0)]
        assert list(out.a) == want, (lo, hi)
        assert out.hr == int(bool(sel) and all(h for _, h, _ in sel))
        assert out.ho == int(bool(sel) and all(o for _, _, o in sel))


def _per_tile_histogram(metas, n, tile):
    """The PerTile loop of allpairs_sum (kernels/storm_allpairs.cpp:547-667)
    driven by the oracle: density order, tile_meta of both tiles,
    select_pairing (Empty -> BB), then per pair the span test (Empty) or the
    tile's cell. Returns the 11 cell_pairs counters."""
    L = O.oracle()
    model = (C.c_double * 16)()
    L.orc_default_model(model)
    order = np.zeros(max(n, 1), np.uint32)
    L.orc_density_order(metas, n, P(order))
    hist = np.zeros(11, np.uint64)
    a, b = OrcMeta(), OrcMeta()
    for i0 in range(0, n, tile):
        i1 = min(i0 + tile, n)
        L.orc_tile_meta(metas, P(order), i0, i1, C.byref(a))
        for j0 in range(i0, n, tile):
            j1 = min(j0 + tile, n)
            L.orc_tile_meta(metas, P(order), j0, j1, C.byref(b))
            tp = L.orc_select_pairing(model, C.byref(a), C.byref(b), None)
            if tp == 10:
                tp = 0
            for i in range(i0, i1):
                mi = metas[int(order[i])].a
                for j in range(i + 1 if j0 == i0 else j0, j1):
                    mj = metas[int(order[j])].a
                    if mi[0] == 0 or mj[0] == 0 or mi[5] < mj[4] or mj[5] < mi[4]:
                        hist[10] += 1
                    else:
                        hist[tp] += 1
    return hist


def _tie_heavy_corpus(u, seed):
    """Rows with many equal cardinalities but different shapes (runs, spans,
    index flags), so which tile a tied row lands in changes the tile means."""
    rng = np.random.default_rng(seed)
    rows = []
    for k in range(120):
        card = int(rng.choice([0, 1, 5, 64, 512, u // 3, u // 2 + 1]))
        if card == 0:
            rows.append(np.zeros(0, np.uint32))
        elif rng.random() < 0.5:            # one run
            s = int(rng.integers(0, u - card + 1))
            rows.append(np.arange(s, s + card, dtype=np.uint32))
        else:                               # scattered
            rows.append(np.sort(rng.choice(u, card, replace=False)).astype(np.uint32))
    offs = np.zeros(len(rows) + 1, np.uint64)
    offs[1:] = np.cumsum([len(r) for r in rows])
    return offs, np.concatenate(rows).astype(np.uint32)


@needs_ref
@pytest.mark.parametrize("shape", ["sweep", "ties"])
def test_tile_meta_and_density_order_reproduce_reference_histogram(shape):
    """The reference's PerTile cell_pairs histogram (allpairs_sum, policy
    PerTile, kernels/storm_allpairs.cpp:547-667) equals the one the oracle's
    density_order + tile_meta + select_pairing + span test produce, over the
    tests/test_cells.cpp:237-287 corpus sweep and a tie-heavy corpus, at tile
    widths 8, 13, 16 and 64. A wrong mean, a lost min/max, an unstable sort or
    a flag OR instead of AND moves pairs between cells."""
    R = O.reference()
    cases = []
    if shape == "sweep":
        for u in (1024, 4033, 16384):
            for structure in (0, 1, 2):
                for spectrum in (0, 1, 2):
                    cases.append((u, *O.ref_generate(40, u, 0.02, structure, spectrum,
                                                      0x9911 + u)))
        cases.append((65536, *O.ref_generate(200, 65536, 0.01, 2, 1, 0x1234)))
    else:
        for u, seed in ((4096, 1), (100000, 2), (1 << 20, 3)):
            cases.append((u, *_tie_heavy_corpus(u, seed)))
    moved = 0
    for u, offs, pos in cases:
        n = len(offs) - 1
        metas = _orc_metas(offs, pos, u)
        o32 = offs.astype(np.uint32)
        for tile in (8, 13, 16, 64):
            want = np.zeros(11, np.uint64)
            R.ref_allpairs_sum(P(o32), P(pos), n, u, 2, tile, P(want), None, None)
            got = _per_tile_histogram(metas, n, tile)
            assert np.array_equal(got, want), (u, n, tile, got, want)
            moved += int((want[:10] > 0).sum() > 1)
    # The sweep must exercise more than one cell per histogram somewhere.
    assert moved > 0


def test_threaded_closed_forms_and_band_agree_with_brute_force():
    """The threaded column closed forms and the triangle band (used on the
    full-size dense configuration) equal the single-thread forms and the
    brute-force triangle on ragged shapes."""
    rng = np.random.default_rng(31)
    for n, nw, p in ((1, 3, 0.5), (2, 1, 0.5), (37, 5, 0.3), (130, 17, 0.05), (65, 64, 0.9)):
        words = np.zeros((n, nw), np.uint64)
        bits = rng.random((n, nw * 64)) < p
        for k in range(64):
            words |= bits[:, k::64].astype(np.uint64) << np.uint64(k)
        t1, rs1 = O.column_closed_forms(words)
        t8, rs8 = O.column_closed_forms(words, threads=7)
        want_sum, tri = O.allpairs_words(words)
        assert t1 == t8 == want_sum
        assert np.array_equal(rs1, rs8)
        for r0, r1 in ((0, 1), (0, n), (n // 3, n // 2 + 1), (n - 1, n)):
            s, band = O.band_words(words, r0, r1, threads=5)
            lo = r0 * n - r0 * (r0 + 1) // 2
            assert np.array_equal(band, tri[lo:lo + len(band)]), (n, r0, r1)
            assert s == int(band.sum(dtype=np.uint64))
