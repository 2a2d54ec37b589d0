"""Host generators of the benchmark workloads (no device needed)."""
import numpy as np


def _valid(offs, pos, u):
    for r in range(len(offs) - 1):
        row = pos[offs[r]:offs[r + 1]].astype(np.int64)
        assert np.all(np.diff(row) > 0) and (len(row) == 0 or (row[0] >= 0 and row[-1] < u))


def test_rows_are_independent_streams(sh):
    o1, p1 = sh.gen_corpus("runs", 64, 1 << 20, 5.0, 100.0, 0.05, 9)
    o2, p2 = sh.gen_corpus("runs", 16, 1 << 20, 5.0, 100.0, 0.05, 9)
    assert np.array_equal(o1[:17], o2) and np.array_equal(p1[:o2[-1]], p2)
    o3, p3 = sh.gen_corpus("runs", 16, 1 << 20, 5.0, 100.0, 0.05, 9)
    assert np.array_equal(o2, o3) and np.array_equal(p2, p3)


def test_zipf_spectrum_shape(sh):
    u = 3866847
    offs, pos = sh.gen_corpus("zipf", 4096, u, 1.1, 386684, 0.0, 3)
    _valid(offs, pos, u)
    card = np.diff(offs.astype(np.int64))
    assert card.min() >= 1 and card.max() <= 386684
    # Zipf(1.1) on [1, 386684]: P(k=1) = 1/H ~ 0.13, mean ~ 1.5e4
    assert 0.11 < np.mean(card == 1) < 0.15
    assert 1.0e4 < card.mean() < 2.0e4


def test_runs_rows_shape(sh):
    u = 36974578
    offs, pos = sh.gen_corpus("runs", 512, u, 20.0, 2000.0, 0.05, 4)
    _valid(offs, pos, u)
    card = np.diff(offs.astype(np.int64))
    assert 3.3e4 < card.mean() < 5.1e4          # ~20 runs x 2000 x 1.05


def test_bernoulli_rows_density(sh):
    offs, pos = sh.gen_corpus("bernoulli", 200, 100000, 0.01, 0, 0, 5)
    _valid(offs, pos, 100000)
    assert abs(offs[-1] / (200 * 100000) - 0.01) < 0.0005
