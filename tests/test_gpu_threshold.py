"""Thresholded mode (storm_hip_allpairs_threshold): exactly the pairs whose
Jaccard reaches t, each with its exact count, against the oracle's per-pair
counts filtered by the reference's test (bench/bench_threshold.cpp jac_ok:
|A n B| > 0 and |A n B| / (|A| + |B| - |A n B|) >= t in double precision)."""
import numpy as np
import pytest

import corpora
import oracle_py as O

pytestmark = pytest.mark.gpu

BUILD_ALL = 4 | 8 | 16
CELLS = ["BB", "BS", "BR", "BW", "SS", "SR", "SW", "RR", "RW", "WW", "EMPTY", "CB"]


def expected(offs, pos, t):
    _, tri = O.allpairs_lists(offs, pos)
    n = len(offs) - 1
    card = np.diff(offs.astype(np.int64)).astype(np.float64)
    I, J = np.triu_indices(n, 1)
    v = tri.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        ok = (tri > 0) & (v / (card[I] + card[J] - v) >= t)
    return np.stack([I[ok], J[ok], tri[ok]], 1).astype(np.uint32)


@pytest.fixture(scope="module")
def cases():
    out = []
    for name, rows, u in corpora.all_corpora():
        if name.startswith("edge") or "-d0.02-" in name or "-d0.3-" in name:
            offs, pos = corpora.csr(rows)
            out.append((name, offs, pos, u))
    return out


@pytest.mark.parametrize("t", [1e-9, 0.01, 0.2, 0.75])
def test_analytic_every_tile_width(ctx, cases, t):
    for name, offs, pos, u in cases:
        want = expected(offs, pos, t)
        print("case", name, "t", t, "hits", len(want), flush=True)
        with ctx.from_lists(offs, pos, u, bitmap_policy=0 | BUILD_ALL) as c:
            for tile in (64, 256):
                st, got = c.allpairs_threshold(t, policy="analytic", tile=tile)
                assert np.array_equal(got, want), (name, t, tile, len(got), len(want))


def test_every_cell_reports_pairs(ctx, cases):
    """Each cell forced on every tile it can run, B x B on both its forms."""
    t = 0.05
    for name, offs, pos, u in cases[:20]:
        want = expected(offs, pos, t)
        print("case", name, "hits", len(want), flush=True)
        with ctx.from_lists(offs, pos, u, bitmap_policy=1 | BUILD_ALL) as c:
            for k, cell in enumerate(CELLS):
                if cell == "EMPTY":
                    continue
                for tile in ((64, 256) if cell == "BB" else (64,)):
                    st, got = c.allpairs_threshold(t, policy="fixed", fixed_cell=k, tile=tile)
                    assert np.array_equal(got, want), (name, cell, tile)


def test_learned_policy_and_graph(ctx, sh):
    offs, pos = sh.gen_corpus("zipf", 1500, 60_000, 1.1, 6_000.0, 0.0, seed=41)
    u = 60_000
    want = expected(offs, pos, 0.02)
    assert 0 < len(want) < 1500 * 1499 // 2
    with ctx.from_lists(offs, pos, u, bitmap_policy=0 | BUILD_ALL) as c:
        for policy in ("analytic", "learned"):
            for use_graph in (False, True, True):
                st, got = c.allpairs_threshold(0.02, policy=policy, tile=256,
                                               use_graph=use_graph)
                assert np.array_equal(got, want), (policy, use_graph)
        # the exact all-pairs call still answers after thresholded ones
        st = c.allpairs(policy="analytic", out="tri", tile=256)
        assert np.array_equal(c.tri(), O.allpairs_lists(offs, pos)[1])


def test_pair_buffer_grows(ctx, sh):
    """More reportable pairs than the first buffer holds (65,536)."""
    offs, pos = sh.gen_corpus("bernoulli", 600, 5_000, 0.1, seed=42)
    want = expected(offs, pos, 0.01)
    assert len(want) > 65_536
    with ctx.from_lists(offs, pos, 5_000, bitmap_policy=0 | BUILD_ALL) as c:
        st, got = c.allpairs_threshold(0.01, policy="analytic", tile=256)
        assert np.array_equal(got, want)


def test_length_filter_prunes_tiles(ctx, sh):
    """Density-sorted tiles whose cardinality ranges cannot reach t are never
    run: with t high on a skewed corpus most tiles are settled as empty."""
    offs, pos = sh.gen_corpus("zipf", 2000, 100_000, 1.1, 10_000.0, 0.0, seed=43)
    with ctx.from_lists(offs, pos, 100_000, bitmap_policy=0 | BUILD_ALL) as c:
        empty = []
        for t in (1e-9, 0.5):
            st, got = c.allpairs_threshold(t, policy="analytic", tile=64)
            assert np.array_equal(got, expected(offs, pos, t)), t
            empty.append(st["cell_tiles"].get("EMPTY", 0))
        assert empty[1] > empty[0], empty


def test_rejects_bad_calls(ctx, sh):
    offs, pos = corpora.csr([np.array([1, 2], np.uint32), np.array([2, 3], np.uint32)])
    with ctx.from_lists(offs, pos, 10, bitmap_policy=BUILD_ALL) as c:
        for t in (0.0, -0.5, 1.5, float("nan")):
            with pytest.raises(sh.StormHipError, match="INVALID|threshold"):
                c.allpairs_threshold(t)
        with pytest.raises(sh.StormHipError):
            c.allpairs_threshold(0.1, policy="all_bitmap")
        st, got = c.allpairs_threshold(1.0)
        assert got.size == 0
        st, got = c.allpairs_threshold(1 / 3)
        assert got.tolist() == [[0, 1, 1]]
    with ctx.gen_bernoulli(300, 5_000, 0.05, 7) as d:
        with pytest.raises(sh.StormHipError):
            d.allpairs_threshold(0.1)


@pytest.mark.parametrize("t", [0.01, 0.2, 0.75, 1.0])
def test_prefix_filter_matches_exact(ctx, cases, t):
    """Prefix-filtered candidates (every mode: forced, gated, exact) report
    exactly the oracle's pairs with their counts."""
    for name, offs, pos, u in cases:
        want = expected(offs, pos, t)
        print("case prefix", name, "t", t, "hits", len(want), flush=True)
        with ctx.from_lists(offs, pos, u, bitmap_policy=0 | BUILD_ALL) as c:
            for mode in ("prefix", "gated", "exact"):
                st, got, info = c.allpairs_threshold_prefix(t, mode=mode)
                assert np.array_equal(got, want), (name, t, mode, len(got), len(want))
                assert info["used_prefix"] == (mode == "prefix" or
                                               (mode == "gated" and info["est_ratio"] < 0.04))
                if info["used_prefix"]:
                    assert info["unique_candidates"] <= info["candidates"]
                    assert st["sum"] == (int(want[:, 2].sum(dtype=np.uint64)) if len(want) else 0)


def prefix_estimate(rows, t):
    """The gate's estimate restated on the CPU: document frequencies, each
    row's prefix of p = |X| - ceil(t|X|) + 1 elements in (floor(log2 df), id)
    order, posting lengths L_e, then (2 sum p + sum_e L_e(L_e-1)/2 mean|X|) /
    (pairs mean|X|) (bench/bench_threshold.cpp:175-196)."""
    allp = np.concatenate(rows).astype(np.int64)
    df = np.bincount(allp)
    b = np.floor(np.log2(np.maximum(df, 1))).astype(np.int64)
    L = np.zeros(len(df), np.int64)
    sump = 0
    for r in rows:
        r = r.astype(np.int64)
        ci = len(r)
        p = min(ci, max(0, ci - int(np.ceil(t * ci)) + 1))
        pre = r[np.argsort(b[r] * (1 << 32) + r, kind="stable")[:p]]
        np.add.at(L, pre, 1)
        sump += p
    n = len(rows)
    mean_s = len(allp) / n
    return (2 * sump + (L * (L - 1) / 2).sum() * mean_s) / (n * (n - 1) / 2 * mean_s)


def test_prefix_filter_sparse_skewed_corpus(ctx):
    """A sparse corpus with skewed element popularity (the regime the prefix
    filter is for): 6,000 rows of 5-40 positions drawn with popularity ~ 1 /
    rank^0.6 over 1,000,000 positions, plus near-duplicate row pairs so that
    high thresholds report pairs. The device's gate estimate equals its CPU
    restatement (which pins the prefixes and postings), the gate picks prefix
    filtering, far fewer candidates than pairs are verified, and the reported
    pairs equal the oracle's at every threshold."""
    rng = np.random.default_rng(31)
    u = 1_000_000
    w = 1.0 / np.arange(1, u + 1) ** 0.6
    w /= w.sum()
    rows = []
    for r in range(6000):
        if r % 10 == 9:
            base = rows[r - 1]
            extra = rng.choice(u, size=2, p=w)
            rows.append(np.unique(np.concatenate([base[: max(1, len(base) - 2)], extra])).astype(np.uint32))
        else:
            rows.append(np.unique(rng.choice(u, size=int(rng.integers(5, 41)), p=w)).astype(np.uint32))
    offs, pos = corpora.csr(rows)
    with ctx.from_lists(offs, pos, u, bitmap_policy=0 | BUILD_ALL) as c:
        for t in (0.1, 0.3, 0.6):
            want = expected(offs, pos, t)
            print("case sparse skewed t", t, "hits", len(want), flush=True)
            st, got, info = c.allpairs_threshold_prefix(t, mode="gated")
            assert abs(info["est_ratio"] - prefix_estimate(rows, t)) <= 1e-9 * info["est_ratio"], info
            assert info["used_prefix"] == 1, info
            assert info["unique_candidates"] < 6000 * 5999 // 2 // 100, info
            assert np.array_equal(got, want), (t, len(got), len(want))


def _skewed_rows(n, u, s, seed):
    rng = np.random.default_rng(seed)
    w = 1.0 / np.arange(1, u + 1) ** s
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    lens = rng.integers(5, 41, size=n)
    draws = np.searchsorted(cdf, rng.random(int(lens.sum())))
    ids = rng.permutation(u)
    offs = np.zeros(n + 1, np.uint64)
    rows, k = [], 0
    for L in lens:
        rows.append(np.unique(ids[draws[k:k + L]]).astype(np.uint32))
        k += L
    offs[1:] = np.cumsum([len(r) for r in rows])
    return offs, np.concatenate(rows).astype(np.uint32)


def test_prefix_filter_many_candidates(ctx):
    """100,000 sparse rows whose popular elements make about 4.4e8 candidate
    emissions (more candidates than one launch has work-items when each takes
    a wavefront): forced prefix filtering reports exactly the pairs and
    counts of the exact thresholded engine, and the gate keeps the exact
    engine here (its estimate is above the cut)."""
    offs, pos = _skewed_rows(100_000, 2_000_000, 0.9, 7)
    with ctx.from_lists(offs, pos, 2_000_000, bitmap_policy=0 | BUILD_ALL) as c:
        _, want, _ = c.allpairs_threshold_prefix(0.1, mode="exact")
        st, got, info = c.allpairs_threshold_prefix(0.1, mode="prefix")
        assert info["candidates"] > (1 << 32) // 64, info
        assert np.array_equal(got, want), (len(got), len(want))
        _, _, ginfo = c.allpairs_threshold_prefix(0.1, mode="gated")
        assert ginfo["used_prefix"] == 0 and ginfo["est_ratio"] >= 0.04, ginfo
