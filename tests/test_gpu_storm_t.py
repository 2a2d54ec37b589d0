"""STORM_t drop-in on the GPU against the compiled reference and the oracle
(tests/test_storm.c:187-270 shapes, plus the square form)."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from storm_t_cases import storm_t_cases
from test_storm_t_host import _fill, gpu_storm

pytestmark = pytest.mark.gpu


def _oracle_sum(vecs):
    uniq = [np.unique(v) for v in vecs]
    offs = np.concatenate([[0], np.cumsum([len(u) for u in uniq])]).astype(np.uint64)
    pos = np.concatenate(uniq).astype(np.uint32) if offs[-1] else np.zeros(0, np.uint32)
    s, _ = O.allpairs_lists(offs, pos, tri=False)
    return s


@pytest.mark.parametrize("k", range(6))
def test_storm_t_pairwise(k):
    label, vecs = storm_t_cases()[k]
    print("case storm_t", label, flush=True)
    G = gpu_storm()
    h = _fill(G, vecs)
    got = G.STORM_pairw_intersect_cardinality(h)
    assert G.STORM_pairw_intersect_cardinality_blocked(h, 7) == got
    assert got == _oracle_sum(vecs), label
    if O.ref_available():
        R = O.reference()
        hr = _fill(R, vecs)
        assert R.STORM_pairw_intersect_cardinality(hr) == got, label
        R.STORM_free(hr)
    G.STORM_free(h)


def test_storm_t_square_and_updates():
    cases = storm_t_cases()
    a, b = cases[0][1][:12], cases[2][1] + cases[1][1][:5]
    G = gpu_storm()
    ha, hb = _fill(G, a), _fill(G, b)
    print("case storm_t square", flush=True)
    got = G.STORM_intersect_cardinality_square(ha, hb)
    want = sum(len(np.intersect1d(x, y)) for x in a for y in b)
    assert got == want
    if O.ref_available():
        R = O.reference()
        ra, rb = _fill(R, a), _fill(R, b)
        assert R.STORM_intersect_cardinality_square(ra, rb) == got
        R.STORM_free(ra)
        R.STORM_free(rb)
    # Adding after a pairwise call re-uploads; clearing empties.
    s0 = G.STORM_pairw_intersect_cardinality(ha)
    extra = a[0]
    assert G.STORM_add(ha, extra.ctypes.data_as(C.c_void_p), len(extra)) == 1
    s1 = G.STORM_pairw_intersect_cardinality(ha)
    assert s1 == s0 + sum(len(np.intersect1d(extra, x)) for x in a)
    assert G.STORM_clear(ha) == 1
    assert G.STORM_pairw_intersect_cardinality(ha) == 0
    G.STORM_free(ha)
    G.STORM_free(hb)


def test_storm_t_duplicates_count_once():
    G = gpu_storm()
    v = np.array([1, 1, 2, 70000, 70000, 70001], dtype=np.uint32)
    w = np.array([1, 2, 2, 70001], dtype=np.uint32)
    h = _fill(G, [v, w])
    assert G.STORM_pairw_intersect_cardinality(h) == 3
    G.STORM_free(h)


def _square_want(a, b):
    ua = [np.unique(x) for x in a]
    ub = [np.unique(y) for y in b]
    return sum(len(np.intersect1d(x, y, assume_unique=True)) for x in ua for y in ub)


def test_storm_t_square_cross_block_and_cache():
    """The square form over the n1 x n2 block only: shapes 1 x N, N x 1 and
    ragged, the same handle on both sides, repeated calls (cached corpus),
    updates to either side, and a freed handle whose address is reused."""
    rng = np.random.default_rng(11)

    def vecs(n, u, k):
        return [np.sort(rng.choice(u, size=int(rng.integers(0, k + 1)), replace=False))
                .astype(np.uint32) for _ in range(n)]

    G = gpu_storm()
    a, b = vecs(1, 300000, 4000), vecs(2500, 300000, 900)
    ha, hb = _fill(G, a), _fill(G, b)
    print("case storm_t square 1 x N", flush=True)
    want = _square_want(a, b)
    assert G.STORM_intersect_cardinality_square(ha, hb) == want
    assert G.STORM_intersect_cardinality_square(ha, hb) == want      # cached
    assert G.STORM_intersect_cardinality_square(hb, ha) == want      # N x 1
    assert G.STORM_intersect_cardinality_square(ha, ha) == _square_want(a, a)
    # An update on either side is seen.
    extra = np.arange(0, 300000, 97, dtype=np.uint32)
    assert G.STORM_add(hb, extra.ctypes.data_as(C.c_void_p), len(extra)) == 1
    b2 = b + [extra]
    assert G.STORM_intersect_cardinality_square(ha, hb) == _square_want(a, b2)
    assert G.STORM_add(ha, extra.ctypes.data_as(C.c_void_p), len(extra)) == 1
    a2 = a + [extra]
    assert G.STORM_intersect_cardinality_square(ha, hb) == _square_want(a2, b2)
    # Free the second side and fill a new handle with other rows.
    G.STORM_free(hb)
    c = vecs(70, 300000, 3000)
    hc = _fill(G, c)
    print("case storm_t square after free", flush=True)
    assert G.STORM_intersect_cardinality_square(ha, hc) == _square_want(a2, c)
    if O.ref_available():
        R = O.reference()
        ra, rc = _fill(R, a2), _fill(R, c)
        assert R.STORM_intersect_cardinality_square(ra, rc) == _square_want(a2, c)
        R.STORM_free(ra)
        R.STORM_free(rc)
    # Ragged: 130 x 77 over a small universe (dense rows, complements).
    d, e = vecs(130, 5000, 4000), vecs(77, 5000, 2500)
    hd, he = _fill(G, d), _fill(G, e)
    assert G.STORM_intersect_cardinality_square(hd, he) == _square_want(d, e)
    for h in (ha, hc, hd, he):
        G.STORM_free(h)
