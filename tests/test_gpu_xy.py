"""Rectangular form (storm_hip_allpairs_xy): |Xi n Yj| for every row of X and
of Y, against the oracle's word kernel per pair, the column closed form
sum_b cX[b] * cY[b] for the total, and the triangle of X . X^T."""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


def _words(rng, n, bits, p):
    nw = (bits + 63) // 64
    b = rng.random((n, nw * 64)) < p
    b[:, bits:] = False
    return np.packbits(b, axis=1, bitorder="little").view(np.uint64).reshape(n, nw)


def _counts(X, Y):
    bx = np.unpackbits(X.view(np.uint8), axis=1, bitorder="little").astype(np.float64)
    by = np.unpackbits(Y.view(np.uint8), axis=1, bitorder="little").astype(np.float64)
    return (bx @ by.T).astype(np.uint32)


@pytest.mark.parametrize("bb", ["auto", "mfma", "mfma_fp4"])
@pytest.mark.parametrize("nx,ny,bits", [(300, 517, 5000), (1, 257, 64), (256, 256, 70001),
                                        (700, 33, 12345)])
def test_rect_matches_oracle(ctx, sh, bb, nx, ny, bits):
    rng = np.random.default_rng(nx * 7 + ny + bits)
    X, Y = _words(rng, nx, bits, 0.07), _words(rng, ny, bits, 0.2)
    print("case", nx, ny, bits, bb, flush=True)
    with ctx.upload_contig(X) as cx, ctx.upload_contig(Y) as cy:
        s, got = ctx.allpairs_xy(cx, cy, bb=bb)
    want = _counts(X, Y)
    assert np.array_equal(got, want)
    assert s == int(want.astype(np.uint64).sum())
    L = O.oracle()
    for i, j in zip(rng.integers(0, nx, 20), rng.integers(0, ny, 20)):
        assert got[i, j] == L.orc_intersect_words(O._p(X[i].copy()), O._p(Y[j].copy()), X.shape[1])


def test_rect_of_a_corpus_with_itself_is_the_triangle(ctx, sh):
    with ctx.gen_bernoulli(900, 20_000, 0.05, 3) as c:
        st = c.allpairs(out="tri", bb="auto")
        tri = c.tri()
        s, M = ctx.allpairs_xy(c, c)
        words = c.download_words(0, 900, (20_000 + 63) // 64)
    iu = np.triu_indices(900, 1)
    assert np.array_equal(M[iu], tri)
    assert np.array_equal(M, M.T)
    card = np.array([int(np.unpackbits(w.view(np.uint8)).sum()) for w in words])
    assert np.array_equal(np.diag(M), card)
    assert s == 2 * st["sum"] + int(card.sum())


def test_rect_large_sum_closed_form(ctx, sh):
    """K-split and many tiles: 3,000 x 2,000 rows of 262,144 bits, the total
    against sum_b cX[b] * cY[b] and sampled pairs against the oracle."""
    bits = 262_144
    with ctx.gen_bernoulli(3000, bits, 0.01, 11) as cx, ctx.gen_bernoulli(2000, bits, 0.03, 12) as cy:
        s, _ = ctx.allpairs_xy(cx, cy, matrix=False)
        s2, M = ctx.allpairs_xy(cx, cy)
        X = cx.download_words(0, 3000, bits // 64)
        Y = cy.download_words(0, 2000, bits // 64)
    assert s == s2 == int(M.astype(np.uint64).sum())
    col = lambda W: np.unpackbits(W.view(np.uint8), axis=1, bitorder="little").sum(0, dtype=np.int64)
    assert s == int((col(X) * col(Y)).sum())
    L = O.oracle()
    rng = np.random.default_rng(5)
    for i, j in zip(rng.integers(0, 3000, 50), rng.integers(0, 2000, 50)):
        assert M[i, j] == L.orc_intersect_words(O._p(X[i].copy()), O._p(Y[j].copy()), bits // 64)


def test_rect_rejects(ctx, sh):
    with ctx.gen_bernoulli(10, 1000, 0.1, 1) as a, ctx.gen_bernoulli(10, 2000, 0.1, 2) as b:
        with pytest.raises(sh.StormHipError):
            ctx.allpairs_xy(a, b)
        with pytest.raises(sh.StormHipError):
            ctx.allpairs_xy(a, a, bb="valu")
    offs = np.array([0, 2, 3], np.uint64)
    pos = np.array([1, 5, 5], np.uint32)
    with ctx.from_lists(offs, pos, 1000) as g, ctx.gen_bernoulli(10, 1000, 0.1, 1) as a:
        with pytest.raises(sh.StormHipError):
            ctx.allpairs_xy(g, a)
