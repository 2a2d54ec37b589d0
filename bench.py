#!/usr/bin/env python3
"""Exact all-pairs intersection cardinality on one MI355X.

Default workload (the headline configuration that fits one GPU): dense
contiguous corpus, N = 32,768 rows x M = 1,048,576 bits (4 GiB packed u64),
every bit i.i.d. Bernoulli(0.05) from the splitmix64 stream Rng(2), generated
on the device; one step = every pair i < j, exact |Xi n Xj| written as a packed
upper-triangle u32 (536,854,528 counts, 2 GiB), inputs resident in HBM.

Prints ONE JSON line on stdout (progress lines go to stderr). `value` = exact
pairs/s of the whole step. `roofline` prices the dominant kernel (B x B) from
HIP events recorded on the engine's own stream around its launch.
`cpu_baseline` times the reference's own all-bitmap B x B kernel (compiled
from its sources, oracle/_ref) on the host cores over a bounded sample of the
same rows.

Other workloads (--config): dense1 (N=2,048 x M=262,144, p=0.01), zipf3 and
runs4 (mixed representations: analytic vs calibrated vs learned selection vs
all-bitmap), batch5 (many independent corpora in one set of launches); their
per-pair output is the upper triangle over density-order positions by default
(--out; the learned call is also timed with the other triangle order). The
default run also reports the thresholded mode on sparse skewed corpora
(prefix-filtered candidates vs the exact thresholded engine).
"""
import argparse
import importlib.util
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "stormbitmaps-internal_amd")
METRIC = "exact pairs/sec on 1 MI355X, learned vs analytic selection; % of roofline"
CELL_NAMES = ["BB", "BS", "BR", "BW", "SS", "SR", "SW", "RR", "RW", "WW", "EMPTY", "CB"]
# Per-pair outputs of the sparse and batched workloads.
OUTPUTS = {
    "tri_density": "u32 upper triangle over density-order positions (the coordinates of "
                   "storm::allpairs_tiles under PerTile; corpus.order() maps them to rows)",
    "tri": "u32 upper triangle in original row order",
    "sum": "u64 sum only",
}

CONFIGS = {
    # name: (n_rows, n_bits, p, seed, description)
    "dense2": (32768, 1 << 20, 0.05, 2,
               "dense contiguous XX^T: N=32,768 x M=1,048,576 bits, Bernoulli(0.05), seed 2, "
               "per-pair u32 upper triangle"),
    "dense1": (2048, 262144, 0.01, 1,
               "dense contiguous XX^T: N=2,048 x M=262,144 bits, Bernoulli(0.01), seed 1, "
               "per-pair u32 upper triangle"),
}
SPARSE_CONFIGS = {
    # name: (generator kind, n_rows, universe, p1, p2, p3, seed, description)
    "zipf3": ("zipf", 65536, 3866847, 1.1, 386684, 0.0, 3,
              "skewed mixed-density rows: N=65,536 over U=3,866,847, cardinality "
              "Zipf(1.1) on [1, 386,684], uniform positions, seed 3"),
    "runs4": ("runs", 16384, 36974578, 20.0, 2000.0, 0.05, 4,
              "run-heavy rows: N=16,384 over U=36,974,578, Poisson(20) runs of geometric "
              "length (mean 2,000) + 5% singletons, seed 4"),
}
BATCH_CONFIGS = {
    "batch5": (512, 5, "batched uneven instances: 512 corpora, N ~ U[64, 4096], universe "
                       "log-uniform [2e5, 1.3e8], density log-uniform [1e-6, 0.17], even "
                       "instances i.i.d. rows, odd instances clustered-run rows, seed 5"),
}
MFMA_I8_PEAK_TOPS = 5000.0   # MI355X dense int8, TOP/s
MFMA_FP4_PEAK_TOPS = 10100.0  # MI355X dense fp4, TFLOP/s
HBM_PEAK_GBS = 8000.0
T0 = time.perf_counter()


def progress(*msg):
    print(f"[bench {time.perf_counter() - T0:7.1f}s]", *msg, file=sys.stderr, flush=True)


def load_storm_hip():
    spec = importlib.util.spec_from_file_location("storm_hip", os.path.join(PKG, "storm_hip.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def traffic_from_profiles(kernel, workload):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), if committed."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def _oracle_mod():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O  # CPU checker: the reference build lives under oracle/_ref
    return O


def host_info(O):
    return {"cpu_model": O.cpu_model(), "reference_isa_build": O.ref_isa(),
            "host_threads": os.cpu_count()}


def cpu_baseline(sh, corpus, n_bits, target_s=12.0, single_target_s=6.0):
    """The reference on this host's cores, beside the GPU (a reported
    baseline): leg (ii) the reference's all-bitmap B x B kernel
    (kernels/cell_bb.cpp 'dense') with project threading over all pairs of the
    first R rows of the same corpus, its sum checked against the GPU's counts
    of the same pairs; leg (i) the reference as shipped on one thread,
    STORM_contig_pairw_intersect_cardinality (storm.cpp:1232-1240) on the
    first rows of the corpus."""
    O = _oracle_mod()
    if not O.ref_available():
        return None
    import ctypes as C
    R_ = O.reference()
    threads = min(16, os.cpu_count() or 1)
    nw = (n_bits + 63) // 64

    def run(r):
        words = corpus.download_words(0, r, nw)
        secs = C.c_double(0)
        s = R_.ref_bb_dense_allpairs(words.ctypes.data_as(C.c_void_p), r, nw, threads,
                                     C.byref(secs))
        return s, secs.value

    r = 256
    _, t = run(r)
    rate = (r * (r - 1) / 2) / max(t, 1e-6)
    r = int(min(8192, max(256, (2 * rate * target_s) ** 0.5), corpus.n))
    progress(f"cpu baseline (threaded reference kernel): {r} rows")
    s, t = run(r)
    pairs = r * (r - 1) // 2
    # The GPU's counts for the same pairs must agree.
    gpu = 0
    n = corpus.n
    for i in range(r - 1):
        off = i * n - i * (i + 1) // 2
        gpu += int(corpus.tri(off, r - i - 1).sum(dtype=np.uint64))
    out = {"value": pairs / t, "unit": "pairs/s", "cores": threads, "kind": "reference",
           "sample": f"all {pairs} pairs of the first {r} rows of the same corpus, "
                     f"{t:.1f} s, reference bb 'dense' kernel, project threading over "
                     f"{threads} threads",
           "sample_sum": int(s), "sample_sum_matches_gpu": bool(gpu == s)}
    out.update(host_info(O))
    # Leg (i): the reference as shipped, one thread.
    r1 = 128
    words = corpus.download_words(0, r1, nw)
    secs = C.c_double(0)
    R_.ref_time_contig(words.ctypes.data_as(C.c_void_p), r1, nw, n_bits, C.byref(secs))
    rate1 = (r1 * (r1 - 1) / 2) / max(secs.value, 1e-6)
    r1 = int(min(4096, max(128, (2 * rate1 * single_target_s) ** 0.5), corpus.n))
    words = corpus.download_words(0, r1, nw)
    s1 = R_.ref_time_contig(words.ctypes.data_as(C.c_void_p), r1, nw, n_bits, C.byref(secs))
    p1 = r1 * (r1 - 1) // 2
    gpu1 = 0
    for i in range(r1 - 1):
        gpu1 += int(corpus.tri(i * n - i * (i + 1) // 2, r1 - i - 1).sum(dtype=np.uint64))
    out["single_thread_reference_as_shipped"] = {
        "value": p1 / secs.value, "unit": "pairs/s", "cores": 1,
        "sample": f"all {p1} pairs of the first {r1} rows, "
                  f"STORM_contig_pairw_intersect_cardinality, {secs.value:.1f} s",
        "sample_sum_matches_gpu": bool(int(s1) == gpu1)}
    return out


def cpu_baseline_sparse(offs, pos, universe, max_rows=2048, tile=64, target_s=8.0,
                        all_bitmap=True):
    """The reference as shipped on one thread (leg (i)) over a stride sample
    of the corpus's rows (every k-th row, so the sample spans the density
    spectrum): allpairs_sum under PerTile with the model calibrate() fits on
    this host, and under AllBitmap (kernels/storm_allpairs.cpp:486-675),
    rows built untimed; each sample sized from a 128-row probe for about
    target_s seconds; every sum checked against the sample's column closed
    form."""
    O = _oracle_mod()
    if not O.ref_available():
        return None
    import ctypes as C
    n = len(offs) - 1
    R_ = O.reference()

    def sample(rows_wanted):
        k = max(1, n // rows_wanted)
        rows = [pos[int(offs[i]):int(offs[i + 1])] for i in range(0, n, k)][:rows_wanted]
        so = np.zeros(len(rows) + 1, np.uint32)
        so[1:] = np.cumsum([len(r) for r in rows])
        sp = np.concatenate(rows).astype(np.uint32) if so[-1] else np.zeros(1, np.uint32)
        return k, rows, so, sp

    def run(policy, cal, rows_wanted):
        k, rows, so, sp = sample(rows_wanted)
        secs = C.c_double(0)
        got = R_.ref_time_allpairs(so.ctypes.data_as(C.c_void_p), sp.ctypes.data_as(C.c_void_p),
                                   len(rows), universe, policy, tile, cal, C.byref(secs))
        want = column_pair_sum(so.astype(np.int64), sp[:int(so[-1])], universe)
        return k, len(rows), secs.value, int(got) == want

    res = {}
    legs = (("per_tile_calibrated", 2, 1), ("all_bitmap", 0, 0))
    for name, policy, cal in legs if all_bitmap else legs[:1]:
        _, r0, t0, _ = run(policy, cal, 128)
        rate = (r0 * (r0 - 1) / 2) / max(t0, 1e-6)
        want_rows = int(min(max_rows, max(128, (2 * rate * target_s) ** 0.5)))
        k, r, t, ok = run(policy, cal, want_rows)
        pairs = r * (r - 1) // 2
        res[name] = {"value": pairs / t, "unit": "pairs/s", "seconds": t, "rows": r,
                     "stride": k, "exact": ok}
        progress(f"cpu reference {name}: {t:.1f} s on {r} rows")
    r = res["per_tile_calibrated"]
    out = {"value": r["value"], "unit": "pairs/s", "cores": 1, "kind": "reference",
           "sample": f"all pairs of every {r['stride']}-th row ({r['rows']} rows), reference "
                     f"allpairs_sum PerTile + calibrate(), one thread, tile {tile}",
           "legs": res,
           "threaded_leg": "not run for this workload (the threaded leg runs the reference's "
                           "B x B kernel on the dense workload)"}
    out.update(host_info(O))
    return out


def column_pair_sum(offs, pos, universe):
    """Sum over pairs of |Xi n Xj| = sum over elements b of C(c_b, 2), c_b =
    number of rows holding b: an exact check independent of every kernel."""
    c = np.bincount(pos, minlength=universe).astype(np.uint64)
    return int((c * (c - np.uint64(1)) // np.uint64(2)).sum(dtype=np.uint64)) if len(c) else 0


def selection_regret(ctx, X, Y, M, idx, decisions):
    """Per-tile regret on held-out sampled tiles: the measured CU-time of each
    policy's chosen cell, summed, over the sum of the fastest measured cell of
    every tile (1.0 = always the exhaustively timed best). `learned_argmin` is
    the network's own ranking before the trust region; the policies' figures
    are the cells their production calls chose."""
    if len(X) == 0:
        return None
    t = np.where(M > 0, np.exp2(Y.astype(np.float64)), np.inf)
    best = t.min(axis=1)
    out = {}
    pred = np.where(M > 0, ctx.selector_predict(X), np.inf)
    choice = pred.argmin(axis=1)
    out["learned_argmin"] = float(t[np.arange(len(t)), choice].sum() / best.sum())
    for name, dec in decisions.items():
        c = dec[idx].astype(np.int64)
        ok = (c < 10)
        ok[ok] &= M[np.nonzero(ok)[0], c[ok]] > 0
        if ok.any():
            out[name] = float(t[np.nonzero(ok)[0], c[ok]].sum() / best[ok].sum())
            out[name + "_tiles"] = int(ok.sum())
    return out


def sparse_bench(sh, args):
    """Mixed-representation workloads: all-bitmap vs analytic vs learned
    per-tile selection, every policy's sum checked equal."""
    kind, n, u, p1, p2, p3, seed, desc = SPARSE_CONFIGS[args.config]
    if args.rows:
        n = args.rows
        desc += f" (first {n} rows)"
    t0 = time.perf_counter()
    offs, pos = sh.gen_corpus(kind, n, u, p1, p2, p3, seed)
    t_gen = time.perf_counter() - t0
    progress(f"generated {n} rows, {len(pos)} elements in {t_gen:.1f} s")
    want = column_pair_sum(offs, pos, u)
    ctx = sh.Context(0)
    # Every representation resident where HBM allows: bitmaps for every row
    # (the all-bitmap store), runs, EWAH and complements; the selectors then
    # choose among all cells per tile. Over the budget, bitmaps for dense rows.
    bitmap_bytes = n * ((u + 63) // 64) * 8
    all_bitmaps = bitmap_bytes <= args.bitmap_budget_gib << 30
    flags = (1 if all_bitmaps else 0) | 4 | 8 | 16
    # Bit-sliced copies of the bitmap groups (as many bytes again) when they fit too.
    if not all_bitmaps or 2 * bitmap_bytes <= args.bitmap_budget_gib << 30:
        flags |= 64
    t0 = time.perf_counter()
    corpus = ctx.from_lists(offs, pos, u, bitmap_policy=flags)
    t_up = time.perf_counter() - t0
    progress(f"built and uploaded in {t_up:.1f} s (bitmaps for every row: {all_bitmaps})")
    build = None if args.no_build_timing else build_timing(sh, ctx, offs, pos, u)
    pos_keep = pos if not args.no_cpu else None
    del pos
    pairs = n * (n - 1) // 2
    res = {}

    def timed(policy, out=None, **kw):
        out = out or args.out
        progress(f"policy {policy}: warm-up")
        t1 = time.perf_counter()
        st = corpus.allpairs(policy=policy, out=out, tile=args.tile, **kw)
        ctx.sync()
        warm = time.perf_counter() - t1
        ts = []
        # A policy slower than --slow-s per call is timed on its warm-up call
        # alone (graph capture included), and says so.
        for k in range(args.steps if warm < args.slow_s else 0):
            t1 = time.perf_counter()
            st = corpus.allpairs(policy=policy, out=out, tile=args.tile, **kw)
            ts.append(time.perf_counter() - t1)
            progress(f"policy {policy}: step {k} {1000 * ts[-1]:.1f} ms")
        single = not ts
        ms = 1000 * float(np.median(ts if ts else [warm]))
        return {"pairs_per_s": pairs / (ms / 1000), "ms": ms, "sum": st["sum"],
                "timed_calls": len(ts) if ts else 1, "single_untimed_warmup": single,
                "cell_tiles": st["cell_tiles"], "skipped": st["skipped"],
                "fallback_tiles": st["fallback_tiles"], "ms_select": st["ms_select"]}

    T = (n + args.tile - 1) // args.tile
    decisions = {}
    if args.only_policy:
        # One policy alone (for profiling its kernels without the others').
        name = args.only_policy
        if name.startswith("fixed_"):
            # Every tile pair forced to one cell (where the tile pair can run it).
            r = timed("fixed", fixed_cell=CELL_NAMES.index(name[6:].upper()), use_graph=True)
        else:
            r = timed("analytic" if name == "analytic_reference_rates" else name,
                      use_graph=name != "all_bitmap")
        out = {"metric": METRIC, "value": r["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1,
               "steps": args.steps, "warmup": 1, "ms_per_step": r["ms"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
               "data": "synthetic (seeded host generator)",
               "config": {"workload": desc, "n_rows": n, "universe": u, "pairs": pairs,
                          "tile": args.tile, "output": OUTPUTS[args.out]},
               "policy": name, "exact": r["sum"] == want, "result": r}
        corpus.free()
        ctx.close()
        return out
    res["analytic_reference_rates"] = timed("analytic", use_graph=True)
    decisions["analytic_reference_rates"] = corpus.last_decisions(T * (T + 1) // 2)
    # One pass of training data feeds both the calibrated analytic model and
    # the learned selector: exploration calls (production calls in which
    # every tile runs a seeded random candidate cell, each item timed in the
    # production mix), or with --sampler sample the per-class timing
    # launches of selector_sample.
    t0 = time.perf_counter()
    parts = []
    chunks = max(1, args.sample_chunks)
    n_hold = 0
    if args.sampler == "explore":
        for ch in range(args.explore_calls):
            X_, Y_, M_, idx_, s_ = ctx.selector_explore(corpus, tile=args.tile, seed=101 + ch,
                                                        fraction=args.explore_fraction)
            if s_ != want:
                raise RuntimeError(f"exploration call {ch}: sum {s_} != closed form {want}")
            parts.append((X_, Y_, M_, idx_))
            progress(f"exploration call {ch + 1}/{args.explore_calls}: {len(X_)} tiles labelled")
    else:
        for ch in range(chunks):
            parts.append(ctx.selector_sample(corpus, tile=args.tile,
                                             max_tiles=max(1, args.sample_tiles // chunks),
                                             seed=11 + ch, repeats=2))
            progress(f"sampled chunk {ch + 1}/{chunks}: {len(parts[-1][0])} tiles")
        # The last chunk is held out of training to measure regret.
        n_hold = len(parts[-1][0]) if chunks > 1 else 0
    X, Y, M, idx = (np.concatenate([p[k] for p in parts]) for k in range(4))
    t_sample = time.perf_counter() - t0
    t0 = time.perf_counter()
    model = ctx.calibrate_samples(corpus, args.tile, Y, M, idx)
    t_cal = time.perf_counter() - t0
    res["analytic"] = timed("analytic", use_graph=True)
    decisions["analytic"] = corpus.last_decisions(T * (T + 1) // 2)
    t0 = time.perf_counter()
    n_fit = len(X) - n_hold
    progress(f"training on {n_fit} tiles ({n_hold} held out)")
    # About 40,000 minibatch steps whatever the number of labelled tiles.
    epochs = max(3, min(300, int(40000 * 128 / max(1, n_fit))))
    loss = ctx.selector_train(X[:n_fit], Y[:n_fit], M[:n_fit], epochs=epochs, batch=128, lr=3e-3,
                              seed=5)
    t_train = time.perf_counter() - t0
    res["learned"] = timed("learned", use_graph=True)
    decisions["learned"] = corpus.last_decisions(T * (T + 1) // 2)
    # The same learned call writing the triangle in original row order
    # (scattered 4-byte writes), for comparison.
    other_out = "tri" if args.out != "tri" else "tri_density"
    learned_other = timed("learned", out=other_out, use_graph=True)
    regret = selection_regret(ctx, X[n_fit:], Y[n_fit:], M[n_fit:], idx[n_fit:], decisions)
    if args.dump_samples:
        np.savez(args.dump_samples, X=X, Y=Y, M=M, idx=idx, n_fit=n_fit,
                 **{"dec_" + k: v for k, v in decisions.items()})
    # Thresholded mode: only pairs with Jaccard >= t, against the exact learned call.
    thresh = {}
    for t in (float(x) for x in args.threshold.split(",") if x):
        ts, n_hits, st = [], 0, None
        for k in range(1 + args.steps):
            t1 = time.perf_counter()
            st, hits = corpus.allpairs_threshold(t, policy="learned", tile=args.tile,
                                                 use_graph=True)
            if k:
                ts.append(time.perf_counter() - t1)
            n_hits = len(hits)
        ms = 1000 * float(np.median(ts))
        thresh[str(t)] = {"pairs_reported": n_hits, "ms": ms, "ms_device": st["ms_device"],
                          "tiles_settled_empty": st["cell_tiles"].get("EMPTY", 0),
                          "over_exact_learned": res["learned"]["ms"] / ms}
        progress(f"threshold {t}: {n_hits} pairs in {ms:.1f} ms")
    if not args.no_allbitmap:
        res["all_bitmap"] = timed("all_bitmap")
    sums = {v["sum"] for v in res.values()}
    out = {
        "metric": METRIC,
        "value": res["learned"]["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": res["learned"]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded host generator)",
        "config": {"workload": desc, "n_rows": n, "universe": u, "pairs": pairs, "tile": args.tile,
                   "output": OUTPUTS[args.out]},
        "exact_sums_agree": len(sums) == 1 and want in sums,
        "sum": sums.pop() if len(sums) == 1 else None, "column_closed_form_sum": want,
        "policies": res,
        "learned_over_analytic": res["learned"]["pairs_per_s"] / res["analytic"]["pairs_per_s"],
        "learned_over_analytic_reference_rates":
            res["learned"]["pairs_per_s"] / res["analytic_reference_rates"]["pairs_per_s"],
        "learned_other_output": {"output": OUTPUTS[other_out], "ms": learned_other["ms"],
                                 "pairs_per_s": learned_other["pairs_per_s"],
                                 "sum_agrees": learned_other["sum"] == want},
        "selector": {"sampler": args.sampler, "labelled_tiles": int(len(X)),
                     "train_loss": loss, "sample_s": t_sample,
                     "train_s": t_train, "held_out_tiles": int(n_hold),
                     "held_out_regret": regret},
        "calibrated_model": model, "calibrate_s": t_cal,
        "setup_s": {"generate": t_gen, "build_upload": t_up},
        "representation_build": build,
        "thresholded": thresh,
    }
    if "all_bitmap" in res:
        out["learned_over_all_bitmap"] = res["learned"]["pairs_per_s"] / res["all_bitmap"]["pairs_per_s"]
    corpus.free()
    ctx.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_sparse(offs, pos_keep, u)
    return out


def skewed_sparse_rows(n, u, s, seed):
    """n rows of 5..40 positions drawn with popularity ~ 1 / rank^s over u
    positions (popular ids scattered by a seeded permutation): a sparse corpus
    with skewed document frequencies, the regime prefix filtering is for."""
    rng = np.random.default_rng(seed)
    w = 1.0 / np.arange(1, u + 1) ** s
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    lens = rng.integers(5, 41, size=n)
    draws = np.searchsorted(cdf, rng.random(int(lens.sum())))
    ids = rng.permutation(u)
    rows, k = [], 0
    for L in lens:
        rows.append(np.unique(ids[draws[k:k + L]]).astype(np.uint32))
        k += L
    offs = np.zeros(n + 1, np.uint64)
    offs[1:] = np.cumsum([len(r) for r in rows])
    return offs, np.concatenate(rows).astype(np.uint32)


def threshold_bench(sh, steps=3):
    """Thresholded mode (pairs with Jaccard >= t, exact counts) on sparse
    skewed corpora: prefix-filtered candidates against the exact thresholded
    engine, the gate's choice, and the pairs each reports (identical sets)."""
    out = []
    ctx = sh.Context(0)
    for n, u, s in ((100_000, 2_000_000, 0.9), (400_000, 4_000_000, 0.6)):
        offs, pos = skewed_sparse_rows(n, u, s, 7)
        with ctx.from_lists(offs, pos, u, bitmap_policy=4 | 8 | 16) as c:
            for t in (0.1, 0.5):
                r = {"rows": n, "universe": u, "popularity_exponent": s, "t": t}
                for mode in ("exact", "prefix", "gated"):
                    ts, res = [], None
                    for k in range(1 + steps):
                        t0 = time.perf_counter()
                        res = c.allpairs_threshold_prefix(t, mode=mode)
                        if k:
                            ts.append(time.perf_counter() - t0)
                    st, hits, info = res
                    r[mode] = {"ms": 1000 * float(np.median(ts)), "pairs_reported": len(hits),
                               "candidates": info["candidates"],
                               "unique_candidates": info["unique_candidates"],
                               "est_ratio": info["est_ratio"],
                               "used_prefix": bool(info["used_prefix"])}
                    if mode == "exact":
                        want = hits
                    else:
                        r[mode]["same_pairs_as_exact"] = bool(np.array_equal(hits, want))
                r["prefix_over_exact"] = r["exact"]["ms"] / r["prefix"]["ms"]
                r["gated_over_exact"] = r["exact"]["ms"] / r["gated"]["ms"]
                progress(f"threshold n={n} t={t}: exact {r['exact']['ms']:.1f} ms, prefix "
                         f"{r['prefix']['ms']:.1f} ms, gate -> "
                         f"{'prefix' if r['gated']['used_prefix'] else 'exact'}")
                out.append(r)
    ctx.close()
    return out


def build_timing(sh, ctx, offs, pos, u):
    """Representation building (runs, EWAH + markers, complements, meta,
    bitmaps of dense rows) three ways: from lists already in HBM on the device,
    from host lists (upload + device build), and on host threads + upload."""
    flags = 0 | 4 | 8 | 16
    to = sh.DeviceArray(np.asarray(offs, dtype=np.uint64))
    tp = sh.DeviceArray(np.asarray(pos, dtype=np.uint32))
    out = {"elements": int(len(pos)), "rows": int(len(offs) - 1), "flags": flags}
    for name, make in (("device_resident_s", lambda: ctx.from_device_lists(to, tp, u, flags)),
                       ("host_lists_device_build_s", lambda: ctx.from_lists(offs, pos, u, flags)),
                       ("host_build_s", lambda: ctx.from_lists(offs, pos, u, flags,
                                                               on_host=True))):
        ts = []
        for _ in range(2):
            t0 = time.perf_counter()
            c = make()
            ts.append(time.perf_counter() - t0)
            c.free()
        out[name] = min(ts)
        progress(f"representation build {name}: {out[name]:.3f} s")
    out["device_elements_per_s"] = len(pos) / out["device_resident_s"]
    out["device_over_host"] = out["host_build_s"] / out["device_resident_s"]
    to.free()
    tp.free()
    return out


def batch_bench(sh, args):
    """The batched workload in full: all 512 independent instances, each
    generated on the device in its stored representation (bitmaps, lists or
    runs), run in HBM-bounded batches of instances; each batch is one set of
    launches per cell class (one captured graph) per policy. Every instance's
    sum is checked against its column closed form (computed on the device
    from the rows themselves, no cell involved); aggregate pairs/s is the
    total pairs over the sum of the batches' call times; per-instance latency
    is the time from its batch's start to its last tile."""
    count, seed, desc = BATCH_CONFIGS[args.config]
    inst = sh.batched_instances(count, seed)
    if args.instances:
        inst = inst[:args.instances]
    budget = args.budget_gib << 30
    batches, cur, used = [], [], 0
    for x in inst:
        if cur and used + x["est_bytes"] > budget:
            batches.append(cur)
            cur, used = [], 0
        cur.append(x)
        used += x["est_bytes"]
    if cur:
        batches.append(cur)
    progress(f"{len(inst)} instances in {len(batches)} batches "
             f"(~{sum(x['est_bytes'] for x in inst) / 2**30:.0f} GiB stored)")
    ctx = sh.Context(0)
    if args.shape:
        ctx.set_bb_config(*(int(x) for x in args.shape.split(",")))
    ref_model = ctx.get_cost_model()
    cal_model = None
    pol_names = ["analytic_reference_rates", "analytic", "learned"]
    res = {k: {"ms": 0.0, "lat": {}, "exact": True, "cell_tiles": {}} for k in pol_names}
    t_gen = t_check = 0.0
    selector = None
    n_done = 0

    def timed(batch, name, wants):
        policy = "learned" if name == "learned" else "analytic"
        t1 = time.perf_counter()
        st, sums, lat = batch.allpairs(policy=policy, out=args.out, tile=args.tile, use_graph=True)
        warm = time.perf_counter() - t1
        progress(f"  {name}: warm-up {1000 * warm:.1f} ms, cells {st['cell_tiles']}")
        ts, lats = [], []
        # A policy slower than --slow-s per call is timed on its warm-up call
        # alone (graph capture included), and says so.
        for k in range(args.steps if warm < args.slow_s else 0):
            t1 = time.perf_counter()
            st, sums, lat = batch.allpairs(policy=policy, out=args.out, tile=args.tile,
                                           use_graph=True)
            ts.append(time.perf_counter() - t1)
            lats.append(lat)
            progress(f"  {name}: step {k} {1000 * ts[-1]:.1f} ms")
        if not ts:
            ts, lats = [warm], [lat]
        ms = 1000 * float(np.median(ts))
        return ms, np.median(np.stack(lats), axis=0), [int(x) for x in sums] == wants, st, \
            len(ts) if warm < args.slow_s else 0

    for bi, group in enumerate(batches):
        corpora, wants = [], []
        t0 = time.perf_counter()
        for x in group:
            corpora.append(sh.gen_instance(ctx, x))
        ctx.sync()
        t_gen += time.perf_counter() - t0
        t0 = time.perf_counter()
        for c in corpora:
            wants.append(int(c.closed_form(row_sums=False)[0]))
        t_check += time.perf_counter() - t0
        dev = sum(c.device_bytes() for c in corpora)
        progress(f"batch {bi + 1}/{len(batches)}: {len(group)} instances generated "
                 f"({dev / 2**30:.1f} GiB on the device)")
        batch = ctx.batch(corpora)
        ctx.put_cost_model(ref_model)
        out = {}
        out["analytic_reference_rates"] = timed(batch, "analytic_reference_rates", wants)
        if selector is None:
            # One sampling pass over instances spread across the size order
            # feeds both the calibrated analytic model (pooled over the
            # sampled instances) and the learned selector.
            t0 = time.perf_counter()
            Xs, Ys, Ms, Is, Cs = [], [], [], [], []
            by_size = sorted(range(len(corpora)), key=lambda i: -corpora[i].n_pairs)
            picked = by_size[::max(1, len(by_size) // 32)][:32]
            for q, i in enumerate(picked):
                X, Y, M, idx = ctx.selector_sample(corpora[i], tile=args.tile,
                                                   max_tiles=max(8, args.sample_tiles // 32),
                                                   seed=11 + i, repeats=2)
                Xs.append(X), Ys.append(Y), Ms.append(M), Is.append(idx)
                Cs.append(np.full(len(X), q, np.uint32))
            X, Y, M = np.concatenate(Xs), np.concatenate(Ys), np.concatenate(Ms)
            t_sample = time.perf_counter() - t0
            progress(f"sampled {len(X)} tiles of {len(picked)} instances in {t_sample:.1f} s")
            ctx.calibrate_corpora([corpora[i] for i in picked], args.tile, Y, M,
                                  np.concatenate(Is), np.concatenate(Cs))
            cal_model = ctx.get_cost_model()
            loss = ctx.selector_train(X, Y, M, epochs=300, batch=128, lr=3e-3, seed=5)
            selector = {"sampled_tiles": int(len(X)), "sampled_instances": len(picked),
                        "train_loss": loss, "sample_s": t_sample}
        ctx.put_cost_model(cal_model)
        out["analytic"] = timed(batch, "analytic", wants)
        out["learned"] = timed(batch, "learned", wants)
        for name, (ms, lat, ok, st, n_timed) in out.items():
            r = res[name]
            r.setdefault("timed_calls", []).append(n_timed)
            r["ms"] += ms
            r["exact"] = r["exact"] and ok
            for x, l in zip(group, lat):
                r["lat"][x["k"]] = float(l)
            for cell, t in st["cell_tiles"].items():
                r["cell_tiles"][cell] = r["cell_tiles"].get(cell, 0) + t
        n_done += len(group)
        batch.free()
        for c in corpora:
            c.free()
    pairs = sum(x["pairs"] for x in inst)
    policies = {}
    for name, r in res.items():
        lat = np.array(list(r["lat"].values()))
        policies[name] = {"pairs_per_s": pairs / (r["ms"] / 1000), "ms": r["ms"],
                          "exact_all_instances": r["exact"],
                          "timed_calls_per_batch": r["timed_calls"],
                          "latency_ms_p50": float(np.percentile(lat, 50)),
                          "latency_ms_p99": float(np.percentile(lat, 99)),
                          "cell_tiles": r["cell_tiles"]}
    out = {
        "metric": METRIC,
        "value": policies["learned"]["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": policies["learned"]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded device generators)",
        "config": {"workload": desc, "instances_run": n_done, "instances_total": count,
                   "batches": len(batches), "pairs": pairs, "tile": args.tile,
                   "output": OUTPUTS[args.out], "batch_budget_gib": args.budget_gib,
                   "instances_with_bitmaps_added": sum(1 for x in inst if x["bitmaps_added"]),
                   "stored_gib_estimate": sum(x["est_bytes"] for x in inst) / 2**30},
        "exact_all": all(p["exact_all_instances"] for p in policies.values()),
        "policies": policies,
        "learned_over_analytic": policies["learned"]["pairs_per_s"] / policies["analytic"]["pairs_per_s"],
        "learned_over_analytic_reference_rates":
            policies["learned"]["pairs_per_s"] / policies["analytic_reference_rates"]["pairs_per_s"],
        "selector": selector, "setup_s": {"generate_on_device": t_gen, "closed_forms": t_check},
        "all_bitmap": "not run: every instance as bitmaps is about 2.7 TB",
    }
    ctx.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_batch(sh, inst)
    return out


def cpu_baseline_batch(sh, inst, per_kind=2, target_s=3.0, max_elements=1 << 26):
    """The reference as shipped on one thread over a bounded sample of the
    batched workload: the first instances of each kind (i.i.d. and run rows)
    with at least 512 rows and at most max_elements positions, built by the
    host generator (the device generator's rows, bit for bit), each timed as
    cpu_baseline_sparse times a corpus (PerTile + calibrate() over stride
    samples of rows, every sum against the sample's closed form;
    the AllBitmap leg is left out here: on these universes a 128-row sample
    alone takes up to 15 s); the aggregate is the sampled pairs over the
    summed seconds."""
    picks = []
    for kind in ("bernoulli", "runs"):
        for x in inst:
            if x["kind"] != kind or x["n"] < 512:
                continue
            if x["n"] * x["universe"] * x["density"] * 1.05 > max_elements:
                continue
            picks.append(x)
            if sum(1 for p in picks if p["kind"] == kind) == per_kind:
                break
    parts = []
    for x in picks:
        kind, n, u, p1, p2, p3, seed, _ = x["gen"]
        offs, pos = sh.gen_corpus(kind, n, u, p1, p2, p3, seed)
        progress(f"cpu reference on instance {x['k']} ({kind}, {n} rows, {len(pos)} positions)")
        r = cpu_baseline_sparse(offs, pos, u, target_s=target_s, all_bitmap=False)
        if r is None:
            return None
        parts.append({"instance": x["k"], "kind": kind, "rows": n, "universe": u,
                      "positions": int(len(pos)), "legs": r["legs"]})
    if not parts:
        return None
    legs = [p["legs"]["per_tile_calibrated"] for p in parts]
    pairs = sum(l["rows"] * (l["rows"] - 1) // 2 for l in legs)
    secs = sum(l["seconds"] for l in legs)
    O = _oracle_mod()
    out = {"value": pairs / secs, "unit": "pairs/s", "cores": 1, "kind": "reference",
           "sample": f"{len(parts)} instances ({', '.join(str(p['instance']) for p in parts)}), "
                     f"stride samples of their rows, reference allpairs_sum PerTile + calibrate(), "
                     f"one thread, tile 64; {pairs} pairs in {secs:.1f} s",
           "exact": all(l["exact"] for p in parts for l in p["legs"].values()),
           "instances": parts}
    out.update(host_info(O))
    return out


def dense_bench(sh, args):
    n, bits, p, seed, desc = CONFIGS[args.config]
    nw = (bits + 63) // 64
    bb = args.bb
    if bb == "auto":
        bb = "mfma_fp4" if nw < (1 << 18) else "mfma"
    ctx = sh.Context(0)
    corpus = ctx.gen_bernoulli(n, bits, p, seed)
    progress(f"generated {n} x {bits} bits on the device; B x B kernel {bb}")
    tuned = None
    if args.shape:
        sw, band, bpc, waves = (int(x) for x in args.shape.split(","))
        ctx.set_bb_config(sw, band, bpc, waves)
        tuned = {"pinned": ctx.get_bb_config()}
    elif bb != "valu" and not args.no_tune:
        # The autotuner picks the launch shape on this corpus (outside the
        # timed steps); the timed steps run with it.
        tuned = ctx.autotune(corpus, bb=bb)
        progress(f"autotuned: {tuned['bb']} {tuned['ms_best']:.1f} ms "
                 f"(starting shape {tuned['ms_default']:.1f} ms, {tuned['n_tried']} shapes)")
    for _ in range(args.warmup):
        corpus.allpairs(out="tri", bb=bb)
    ctx.sync()
    t0 = time.perf_counter()
    bb_ms, dev_ms, sums = [], [], set()
    for k in range(args.steps):
        st = corpus.allpairs(out="tri", bb=bb)
        bb_ms.append(st["ms_bb"])
        dev_ms.append(st["ms_device"])
        sums.add(st["sum"])
        progress(f"step {k}: B x B {st['ms_bb']:.1f} ms")
    ctx.sync()
    wall = time.perf_counter() - t0
    ms_step = wall * 1000.0 / args.steps
    pairs = n * (n - 1) // 2
    kern_ms = float(np.mean(bb_ms))
    if bb in ("mfma", "mfma_fp4"):
        fp4 = bb == "mfma_fp4"
        kernel = "bb_mfma_fp4" if fp4 else "bb_mfma_i8"
        peak = MFMA_FP4_PEAK_TOPS if fp4 else MFMA_I8_PEAK_TOPS
        ops = 2.0 * pairs * bits                      # MACs x 2, exact pairs only
        achieved = ops / (kern_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak,
                "traffic": traffic_from_profiles(kernel, args.config),
                "kernel": kernel, "kernel_ms": kern_ms,
                "algorithmic": "2*M ops per exact pair i<j (0/1 operands, "
                               + ("e2m1 with f32" if fp4 else "int8 with i32") + " accumulate)"}
    else:
        kernel = "bb_valu"
        wp = pairs * nw
        valu_peak = 256 * 64 * 2.4e9 / 4            # 64-bit word-pairs/s (4 VALU ops each)
        achieved = wp / (kern_ms * 1e-3)
        roof = {"bound": "valu", "achieved": achieved / 1e12, "peak": valu_peak / 1e12,
                "unit": "T word-pairs/s", "frac": achieved / valu_peak,
                "traffic": traffic_from_profiles(kernel, args.config),
                "kernel": kernel, "kernel_ms": kern_ms,
                "algorithmic": "ceil(M/64) 64-bit AND+popcount per exact pair i<j"}
    out = {
        "metric": METRIC,
        "value": pairs / (ms_step / 1000.0),
        "unit": "pairs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"mfma": "int8", "mfma_fp4": "fp4 (e2m1 0/1, f32 exact accumulate)"}.get(bb, "u32"),
        "data": "synthetic (device splitmix64 Bernoulli stream)",
        "config": {"workload": desc, "n_rows": n, "n_bits": bits, "density": p, "seed": seed,
                   "pairs": pairs, "bb_kernel": bb},
        "exact_sum": sums.pop() if len(sums) == 1 else None,
        "ms_device_per_step": float(np.mean(dev_ms)),
        "roofline": roof,
        "launch_shape": tuned,
        "cpu_baseline": None if args.no_cpu else cpu_baseline(sh, corpus, bits),
    }
    corpus.free()
    ctx.close()
    return out


def copy_probe(args):
    """A plain device copy of known size (1 GiB read + 1 GiB written per
    launch), for scaling the HBM byte counters a profiler reads on the B x B
    kernel. Not a bench line: prints the bytes it moved."""
    import torch
    n = 1 << 28                                    # 2^28 int32 = 1 GiB
    a = torch.ones(n, dtype=torch.int32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(max(1, args.steps)):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read_per_launch": 4 * n, "copy_bytes_written_per_launch": 4 * n,
                      "launches": max(1, args.steps)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="dense2",
                    choices=sorted(CONFIGS) + sorted(SPARSE_CONFIGS) + sorted(BATCH_CONFIGS)
                    + ["copy"])
    ap.add_argument("--bb", default="auto", choices=["auto", "mfma", "mfma_fp4", "valu"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="default run: the headline workload only (no other workloads)")
    ap.add_argument("--no-tune", action="store_true",
                    help="dense configs: keep the default B x B launch shape")
    ap.add_argument("--shape", default="",
                    help="pin the B x B launch shape stage_words,band,blocks_per_cu,waves "
                         "(dense configs: no tuning; batch configs: every call)")
    ap.add_argument("--rows", type=int, default=0, help="sparse configs: first N rows only")
    ap.add_argument("--out", default="tri_density", choices=sorted(OUTPUTS),
                    help="sparse and batch configs: per-pair output (the triangle packed over "
                         "density-order positions, over original rows, or the sum only)")
    ap.add_argument("--tile", type=int, default=256,
                    help="rows per selection tile (sparse and batch configs)")
    ap.add_argument("--slow-s", type=float, default=30.0,
                    help="sparse configs: a policy slower than this per call runs once")
    ap.add_argument("--bitmap-budget-gib", type=int, default=160,
                    help="sparse configs: keep a bitmap of every row when they fit this")
    ap.add_argument("--sample-tiles", type=int, default=1024)
    ap.add_argument("--only-policy", default="",
                    choices=["", "analytic_reference_rates", "all_bitmap", "fixed_bb",
                             "fixed_bs", "fixed_rr", "fixed_ss"],
                    help="sparse configs: time this one policy only (no sampling or training)")
    ap.add_argument("--dump-samples", default="",
                    help="sparse configs: write the selector's sampled tiles (features, "
                         "measured log2 CU-time per cell, masks) and each policy's decisions "
                         "to this .npz")
    ap.add_argument("--sample-chunks", type=int, default=8,
                    help="sampling passes (a progress line after each)")
    ap.add_argument("--sampler", default="explore", choices=["explore", "sample"],
                    help="sparse configs: training data from exploration calls in the "
                         "production mix, or from per-class sampling launches")
    ap.add_argument("--explore-calls", type=int, default=8,
                    help="sparse configs: exploration calls (one label per tile each)")
    ap.add_argument("--explore-fraction", type=float, default=0.125,
                    help="sparse configs: share of tiles that run a random candidate cell in "
                         "each exploration call (the others run the analytic choice)")
    ap.add_argument("--no-allbitmap", action="store_true")
    ap.add_argument("--no-build-timing", action="store_true",
                    help="sparse configs: skip timing the representation builders")
    ap.add_argument("--threshold", default="0.01,0.1,0.5",
                    help="sparse configs: Jaccard thresholds to time in thresholded mode "
                         "(comma list, empty for none)")
    ap.add_argument("--budget-gib", type=int, default=150,
                    help="batch configs: estimated stored size of one batch of instances")
    ap.add_argument("--instances", type=int, default=0,
                    help="batch configs: the first N instances only (0: all)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(the all-pairs triangle would shard by tile rows across GPUs, each holding "
              "every row panel)", file=sys.stderr)
        sys.exit(2)

    if args.config == "copy":
        copy_probe(args)
        return
    sh = load_storm_hip()
    if args.config in SPARSE_CONFIGS:
        out = sparse_bench(sh, args)
    elif args.config in BATCH_CONFIGS:
        out = batch_bench(sh, args)
    else:
        out = dense_bench(sh, args)
    if args.config == "dense2" and not args.no_extra:
        out["other_configs"] = other_configs(sh, args)
        try:
            out["other_configs"]["threshold_sparse"] = threshold_bench(sh)
        except Exception as e:   # the exact-count lines above stand on their own
            out["other_configs"]["threshold_sparse"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


def other_configs(sh, args):
    """The default run's other workloads, each at its stated size, after the
    headline line's own measurement: the small dense workload (fp4 B x B,
    full), the skewed and the run-heavy mixed workloads (analytic with the
    reference's rates, learned trained on two exploration calls, all-bitmap)
    with the reference's single-thread CPU leg on a row sample. The batched
    workload (about five minutes) runs with --config batch5."""
    import copy
    out = {}
    for name in ("dense1", "zipf3", "runs4"):
        a = copy.copy(args)
        a.config = name
        a.steps = 3
        a.warmup = 1
        a.no_build_timing = True
        a.threshold = ""
        a.explore_calls = 2
        a.explore_fraction = 1.0 / 32
        a.slow_s = 5.0
        a.no_tune = False
        a.shape = ""
        t0 = time.perf_counter()
        progress(f"other workload {name}")
        try:
            r = sparse_bench(sh, a) if name in SPARSE_CONFIGS else dense_bench(sh, a)
            keep = ("value", "unit", "ms_per_step", "dtype", "config", "roofline", "exact_sum",
                    "exact_sums_agree", "column_closed_form_sum", "cpu_baseline",
                    "learned_over_analytic_reference_rates", "learned_over_analytic",
                    "learned_over_all_bitmap", "learned_other_output", "selector",
                    "launch_shape")
            r2 = {k: r[k] for k in keep if k in r}
            if "policies" in r:
                r2["policies"] = {k: {q: v[q] for q in ("pairs_per_s", "ms", "cell_tiles",
                                                         "timed_calls") if q in v}
                                  for k, v in r["policies"].items()}
            r2["wall_s"] = time.perf_counter() - t0
            out[name] = r2
        except Exception as e:   # one workload's failure does not hide the others
            out[name] = {"error": repr(e)}
    return out


if __name__ == "__main__":
    main()
