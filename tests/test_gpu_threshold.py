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
