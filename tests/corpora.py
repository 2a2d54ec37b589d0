"""Seeded test corpora as CSR (offs u64[n+1], pos u32 sorted per row).

Shapes follow the reference's test sweeps (tests/test_cells.cpp:195-230 edge
rows; :302-325 structure x spectrum x density x universe)."""
import zlib

import numpy as np


def csr(rows):
    offs = np.zeros(len(rows) + 1, dtype=np.uint64)
    if rows:
        offs[1:] = np.cumsum([len(r) for r in rows])
    pos = (np.concatenate([np.asarray(r, dtype=np.uint32) for r in rows])
           if rows and offs[-1] else np.zeros(0, dtype=np.uint32))
    return offs, pos.astype(np.uint32)


def edge_rows(u):
    rows = [[], [0], [u - 1], [63, 64], [511, 512], list(range(u)), list(range(0, u, 64)),
            list(range(0, u, 2)), list(range(u // 2, min(u // 2 + 300, u))),
            list(range(1, min(200, u)))]
    return [np.array(sorted({p for p in r if p < u}), dtype=np.uint32) for r in rows]


def uniform_row(rng, u, k):
    k = int(min(max(k, 0), u))
    if k == 0:
        return np.zeros(0, np.uint32)
    if k > u // 2:
        keep = np.ones(u, bool)
        keep[rng.choice(u, u - k, replace=False)] = False
        return np.flatnonzero(keep).astype(np.uint32)
    return np.sort(rng.choice(u, k, replace=False)).astype(np.uint32)


def runs_row(rng, u, k, mean_run):
    """Alternating geometric gaps and runs (the gap model of
    kernels/storm_gen.cpp:121-169), stopping at k bits or the universe end."""
    d = min(0.999, k / u)
    gap = max(1.0, mean_run * (1 - d) / d)
    out, got, p = [], 0, 0
    while got < k and p < u:
        p += int(rng.geometric(1.0 / gap))
        if p >= u:
            break
        ln = int(min(rng.geometric(1.0 / max(mean_run, 1.0)), k - got, u - p))
        out.append(np.arange(p, p + ln))
        got += ln
        p += ln
    return (np.concatenate(out) if out else np.zeros(0)).astype(np.uint32)


def mixed_corpus(seed, n, u, density, structure="uniform", spectrum="inverse", mean_run=32):
    rng = np.random.default_rng(seed)
    T = density * u
    rows = []
    for _ in range(n):
        if spectrum == "uniform":
            k = int(round(T))
        elif spectrum == "inverse":
            C = max(2.0, 4 * T)
            k = int(np.floor(np.exp(rng.random() * np.log(C))))
        else:  # bimodal
            k = int(max(1, (T - 0.9) / 0.1)) if rng.random() < 0.1 else 1
        k = max(1, min(u, k))
        if structure == "uniform":
            rows.append(uniform_row(rng, u, k))
        else:
            rows.append(runs_row(rng, u, k, mean_run))
    return rows


def all_corpora():
    """(name, rows, universe) covering every code path the cells branch on."""
    out = []
    for u in (64, 1025, 4033, 70001):
        out.append((f"edge-u{u}", edge_rows(u), u))
    for u in (1024, 4033, 16384, 65536):
        for st in ("uniform", "runs"):
            for sp in ("uniform", "inverse", "bimodal"):
                for d in (5e-4, 0.02, 0.3):
                    seed = zlib.crc32(f"{u}{st}{sp}{d}".encode())
                    out.append((f"{st}-{sp}-d{d}-u{u}",
                                mixed_corpus(seed, 37, u, d, st, sp), u))
    return out
