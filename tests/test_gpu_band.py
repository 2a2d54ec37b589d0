"""Streamed tile rows (storm::allpairs_tiles, kernels/storm_allpairs.cpp:677-732)
and cross blocks (STORM_intersect_cardinality_square, storm.cpp:1027-1046) on
the GPU, against the CPU oracle's per-pair counts."""
import zlib

import numpy as np
import pytest

import corpora
import oracle_py as O

pytestmark = pytest.mark.gpu

BUILD_ALL = 4 | 8 | 16


def _square(rows):
    """Symmetric n x n matrix of |Xi n Xj| from the oracle's triangle."""
    offs, pos = corpora.csr(rows)
    s, tri = O.allpairs_lists(offs, pos)
    n = len(rows)
    m = np.zeros((n, n), dtype=np.uint32)
    iu = np.triu_indices(n, 1)
    m[iu] = tri
    return offs, pos, s, m + m.T


def _corpus(name, n, u, density, structure, spectrum):
    seed = zlib.crc32(f"band{name}".encode())
    return corpora.mixed_corpus(seed, n, u, density, structure, spectrum)


CASES = [
    ("uniform-inverse", 700, 16384, 0.02, "uniform", "inverse"),
    ("runs-bimodal", 600, 65536, 0.02, "runs", "bimodal"),
    ("uniform-dense", 520, 4033, 0.3, "uniform", "uniform"),
    ("runs-sparse", 450, 70001, 5e-4, "runs", "inverse"),
]


@pytest.fixture(scope="module")
def cases():
    out = []
    for name, n, u, d, st, sp in CASES:
        rows = _corpus(name, n, u, d, st, sp)
        out.append((name, u) + _square(rows))
    return out


def _expected_band(m, perm, i0, i1):
    sub = m[np.ix_(perm[i0:i1], perm)]
    mask = np.arange(m.shape[0])[None, :] > np.arange(i0, i1)[:, None]
    return np.where(mask, sub, 0).astype(np.uint32)


@pytest.mark.parametrize("policy", ["analytic", "learned", "all_bitmap"])
def test_bands_cover_the_triangle(ctx, cases, policy):
    """Every band in both orders, every tile height: the counts of the band's
    pairs in the call's order, zeros below the diagonal, band sums adding up
    to the total."""
    for name, u, offs, pos, want_sum, m in cases:
        n = m.shape[0]
        with ctx.from_lists(offs, pos, u, bitmap_policy=BUILD_ALL) as c:
            density = c.order()
            # Stable by cardinality (kernels/storm_allpairs.cpp:316-322).
            card = np.diff(offs)
            assert np.array_equal(density, np.argsort(card, kind="stable")), name
            for tile in (64, 128, 256):
                for order in ("density", "identity"):
                    print("case band", name, policy, tile, order, flush=True)
                    perm = density if order == "density" else np.arange(n)
                    total = 0
                    for band in range((n + tile - 1) // tile):
                        st, got = c.allpairs_band(band, policy=policy, order=order, tile=tile)
                        i0, i1 = band * tile, min(n, band * tile + tile)
                        want = _expected_band(m, perm, i0, i1)
                        assert np.array_equal(got, want), (name, policy, tile, order, band)
                        assert st["sum"] == int(want.sum(dtype=np.uint64)), (name, band)
                        assert st["pairs"] == sum(n - i - 1 for i in range(i0, i1))
                        total += st["sum"]
                    assert total == want_sum, (name, policy, tile, order)


def test_band_fixed_cells_and_graph(ctx, cases):
    """Forced cells and graph replays give the same bands (one graph serves
    every band: the range is read on the device)."""
    name, u, offs, pos, want_sum, m = cases[0]
    n = m.shape[0]
    with ctx.from_lists(offs, pos, u, bitmap_policy=1 | BUILD_ALL) as c:
        perm = c.order()
        for cell in (0, 1, 4, 7, 11):
            for band in (0, 3, (n - 1) // 64):
                print("case band fixed", cell, band, flush=True)
                _, got = c.allpairs_band(band, policy="fixed", fixed_cell=cell, tile=64)
                assert np.array_equal(got, _expected_band(m, perm, band * 64,
                                                          min(n, band * 64 + 64))), (cell, band)
        for rep in range(2):
            for band in range((n + 127) // 128):
                _, got = c.allpairs_band(band, policy="analytic", tile=128, use_graph=True)
                want = _expected_band(m, perm, band * 128, min(n, band * 128 + 128))
                assert np.array_equal(got, want), (rep, band)
        # A whole-triangle call on the same corpus afterwards is unaffected.
        st = c.allpairs(policy="analytic", out="tri", tile=64)
        assert st["sum"] == want_sum
        assert st["pairs"] == n * (n - 1) // 2
        iu = np.triu_indices(n, 1)
        assert np.array_equal(c.tri(), m[iu])


def test_band_rejects(ctx, sh, cases):
    name, u, offs, pos, want_sum, m = cases[0]
    with ctx.from_lists(offs, pos, u, bitmap_policy=BUILD_ALL) as c:
        with pytest.raises(sh.StormHipError):
            c.allpairs_band(99, tile=64)
        with pytest.raises(sh.StormHipError):
            c.allpairs_band(0, tile=96)
        with pytest.raises(sh.StormHipError):
            c.allpairs_band(0, policy="exhaustive")


def _cross_corpus(rows_a, rows_b, tile):
    """Rows of A, empty rows up to a tile multiple, rows of B."""
    pad = (-len(rows_a)) % tile
    rows = list(rows_a) + [np.zeros(0, dtype=np.uint32)] * pad + list(rows_b)
    return rows, len(rows_a) + pad


@pytest.mark.parametrize("tile", [64, 128, 256])
def test_cross_block_every_pair_across(ctx, cases, tile):
    """The n1 x n2 block only, in row order: equal to the oracle's counts,
    the sum-only form agreeing with it."""
    for k, (name, u, offs, pos, want_sum, m) in enumerate(cases):
        n = m.shape[0]
        rows = [pos[offs[r]:offs[r + 1]] for r in range(n)]
        for n1 in (1, 77, n // 2):
            print("case cross", name, tile, n1, flush=True)
            a, b = rows[:n1], rows[n1:]
            allrows, split = _cross_corpus(a, b, tile)
            o2, p2 = corpora.csr(allrows)
            with ctx.from_lists(o2, p2, u, bitmap_policy=BUILD_ALL) as c:
                for policy in ("analytic", "all_bitmap") if k == 0 else ("analytic",):
                    st, got = c.allpairs_cross(split, policy=policy, tile=tile)
                    want = np.zeros((split, n - n1), dtype=np.uint32)
                    want[:n1] = m[:n1, n1:]
                    assert np.array_equal(got, want), (name, tile, n1, policy)
                    assert st["sum"] == int(want.sum(dtype=np.uint64))
                    assert st["pairs"] == split * (n - n1)
                    st2, none = c.allpairs_cross(split, policy=policy, tile=tile, matrix=False)
                    assert none is None and st2["sum"] == st["sum"], (name, tile, n1, policy)


def test_cross_rejects_split_off_the_tile_grid(ctx, sh, cases):
    name, u, offs, pos, want_sum, m = cases[0]
    with ctx.from_lists(offs, pos, u, bitmap_policy=BUILD_ALL) as c:
        for split in (0, 65, m.shape[0]):
            with pytest.raises(sh.StormHipError):
                c.allpairs_cross(split, tile=64)
