"""Host-side output layouts (no device needed)."""
import numpy as np


def _tri_index(i, j, n):
    lo, hi = min(i, j), max(i, j)
    return lo * n - lo * (lo + 1) // 2 + (hi - lo - 1)


def test_density_triangle_rearrangement(sh):
    """out="tri_density" packs the count of rows perm[a], perm[b] at position
    pair (a, b); Corpus.tri_to_original puts every count back at its rows'
    place in the original-order triangle. Checked against a triangle built
    pair by pair from a random permutation (with the identity as a case)."""
    rng = np.random.default_rng(3)

    class Stub:
        pass

    for n, perm in ((2, np.array([1, 0])), (37, rng.permutation(37)), (64, np.arange(64))):
        perm = perm.astype(np.uint32)
        I, J = np.triu_indices(n, 1)
        orig = rng.integers(0, 1 << 31, len(I)).astype(np.uint32)
        dens = np.array([orig[_tri_index(int(perm[a]), int(perm[b]), n)] for a, b in zip(I, J)],
                        dtype=np.uint32)
        stub = Stub()
        stub.n = n
        stub.order = lambda perm=perm: perm
        assert np.array_equal(sh.Corpus.tri_to_original(stub, dens), orig), n
