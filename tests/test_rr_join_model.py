"""The R x R interval join's chunk / bucket / entry-pool logic (the device
kernel rr_tile_join in stormbitmaps-internal_amd/csrc/cells.hip), restated on
the CPU step by step and checked against brute-force run overlaps. Tiny pools
force every path the device takes rarely: entry-pool overflow and the retry
at half width, chunk widening and narrowing, runs straddling chunk and
bucket edges, runs covering whole buckets, empty rows and a one-position
universe. The device kernel itself is checked bit-exact on the GPU
(tests/test_gpu_general.py); this pins the algorithm it implements."""
import random

import numpy as np


def runs_from_bits(bits):
    s = []; i = 0; n = len(bits)
    while i < n:
        if bits[i]:
            j = i
            while j < n and bits[j]: j += 1
            s.append((i, j)); i = j
        else:
            i += 1
    return s

def lower_bound(a, x):
    lo, hi = 0, len(a)
    while lo < hi:
        m = (lo + hi) // 2
        if a[m] < x: lo = m + 1
        else: hi = m
    return lo

def join(X, Y, U, NB, NE, MAXSH):
    nx, ny = len(X), len(Y)
    cnt = [[0] * 64 for _ in range(64)]
    total = sum(len(r) for r in X)
    want = U * (NE // 2) // total if total else U
    sh = 0
    while sh < MAXSH and (NB << (sh + 1)) <= want: sh += 1
    st = [[0] * 64 for _ in range(2)]
    en = [[0] * 64 for _ in range(2)]
    lo = 0
    chunks = retries = 0
    while lo < U:
        hi = min(U, lo + (NB << sh))
        for side, rows in ((0, X), (1, Y)):
            for r, R in enumerate(rows):
                ss = [a for a, b in R]
                en[side][r] = len(R) if hi >= U else st[side][r] + lower_bound(ss[st[side][r]:], hi)
        w = 1 << sh
        cover = [0] * NB; heads = [[] for _ in range(NB)]; n_ent = 0
        for r, R in enumerate(X):
            for k in range(st[0][r], en[0][r]):
                s, e = R[k]
                rs = max(s, lo) - lo; re = min(e, hi) - lo
                assert rs < re
                kb = rs >> sh; ke = (re - 1) >> sh
                assert ke < NB
                def part(b, f, l):
                    nonlocal n_ent
                    x = n_ent; n_ent += 1
                    if x >= NE: return
                    assert 0 <= f <= l < w and l < 65536
                    heads[b].append((r, f, l))
                if kb == ke:
                    if re - rs == w: cover[kb] |= 1 << r
                    else: part(kb, rs & (w - 1), (re - 1) & (w - 1))
                else:
                    if rs & (w - 1): part(kb, rs & (w - 1), w - 1)
                    else: cover[kb] |= 1 << r
                    for b in range(kb + 1, ke): cover[b] |= 1 << r
                    if re & (w - 1): part(ke, 0, (re - 1) & (w - 1))
                    else: cover[ke] |= 1 << r
        if n_ent > NE:
            assert sh > 0
            sh -= 1; retries += 1
            continue
        for r, R in enumerate(Y):
            for k in range(st[1][r], en[1][r]):
                s, e = R[k]
                rs = max(s, lo) - lo; re = min(e, hi) - lo
                kb = rs >> sh; ke = (re - 1) >> sh
                for b in range(kb, ke + 1):
                    b0 = b << sh
                    qs = max(rs, b0) - b0; qe = min(re, b0 + w) - b0
                    m = cover[b]
                    a = 0
                    while m:
                        if m & 1: cnt[a][r] += qe - qs
                        m >>= 1; a += 1
                    for (a, es, el) in heads[b]:
                        o0 = max(qs, es); o1 = min(qe, el + 1)
                        if o0 < o1: cnt[a][r] += o1 - o0
        for side, rows in ((0, X), (1, Y)):
            for r, R in enumerate(rows):
                e_ = en[side][r]; s_ = st[side][r]
                st[side][r] = e_ - 1 if (e_ > s_ and R[e_ - 1][1] > hi) else e_
        lo = hi
        chunks += 1
        if n_ent < NE // 8 and sh < MAXSH: sh += 1
        elif n_ent > NE * 3 // 4 and sh > 0: sh -= 1
    return cnt, chunks, retries

def brute(X, Y):
    """Pairwise overlap of run rows via dense 0/1 rows (independent of runs)."""
    U = 1 + max([e for R in X + Y for _, e in R] + [0])
    def dense(R):
        v = np.zeros(U, np.int64)
        for s_, e_ in R:
            v[s_:e_] = 1
        return v
    A = np.array([dense(R) for R in X]).reshape(len(X), U)
    B = np.array([dense(R) for R in Y]).reshape(len(Y), U)
    C = A @ B.T
    out = [[0] * 64 for _ in range(64)]
    for a in range(len(X)):
        for b in range(len(Y)):
            out[a][b] = int(C[a, b])
    return out

def gen_row(rng, U, kind):
    bits = [0] * U
    if kind == 0:
        for _ in range(rng.randint(0, 6)):
            s = rng.randrange(U); L = rng.choice([1, 2, 5, 40, 300, U])
            for p in range(s, min(U, s + L)): bits[p] = 1
        for _ in range(rng.randint(0, 30)): bits[rng.randrange(U)] = 1
    elif kind == 1:
        p = rng.random() * 0.5
        for i in range(U): bits[i] = 1 if rng.random() < p else 0
    else:
        pass
    return runs_from_bits(bits)



def test_join_model_matches_brute_force():
    rng = random.Random(7)
    retries = 0
    for trial in range(60):
        U = rng.choice([1, 7, 64, 500, 3000, 20000])
        nx, ny = rng.randint(1, 12), rng.randint(1, 12)
        X = [gen_row(rng, U, rng.choice([0, 0, 1, 2])) for _ in range(nx)]
        Y = [gen_row(rng, U, rng.choice([0, 0, 1, 2])) for _ in range(ny)]
        NB, NE, MAXSH = rng.choice([(4, 8, 3), (16, 32, 6), (64, 128, 16)])
        got, _, rt = join(X, Y, U, NB, NE, MAXSH)
        retries += rt
        assert got == brute(X, Y), (trial, U, NB, NE)
    assert retries > 0   # the overflow path ran
