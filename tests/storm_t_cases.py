"""Vector sets for the STORM_t drop-in tests, in the shape of
tests/test_storm.c:187-270 (xorshift32 from seed 12345, make_set)."""
import numpy as np


class Xorshift32:
    """tests/test_storm.c:56 xorshift32."""

    def __init__(self, s=12345):
        self.s = s

    def __call__(self):
        x = self.s
        x ^= (x << 13) & 0xFFFFFFFF
        x ^= x >> 17
        x ^= (x << 5) & 0xFFFFFFFF
        self.s = x
        return x


def make_set(rng, n, universe, dups=False):
    v = sorted(rng() % universe for _ in range(n))
    if not dups:
        v = sorted(set(v))
    return np.array(v, dtype=np.uint32)


def storm_t_cases():
    """(label, list of vectors) -- multi-block sparse, very sparse, one block
    with dense (bitmap) and sparse (scalar) vectors alternating, duplicates,
    empty vectors, values near 2^32."""
    rng = Xorshift32()
    out = []
    out.append(("multi-block sparse",
                [make_set(rng, 1 + rng() % 299, 200000) for _ in range(30)]))
    out.append(("very sparse", [make_set(rng, 1 + rng() % 2, 100000) for _ in range(30)]))
    out.append(("mixed dense/scalar",
                [make_set(rng, 9000 if i % 2 == 0 else 50, 65536) for i in range(12)]))
    out.append(("many blocks", [make_set(rng, 20000, 3_000_000) for _ in range(20)]))
    out.append(("with empty", [make_set(rng, 100, 5000) for _ in range(5)]
                + [np.zeros(0, np.uint32)] + [make_set(rng, 100, 5000) for _ in range(5)]))
    hi = 0xFFFFFFFE
    out.append(("high values", [np.array(sorted({hi - rng() % 70000 for _ in range(3000)}),
                                         dtype=np.uint32) for _ in range(8)]))
    return out
