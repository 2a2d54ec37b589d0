#!/usr/bin/env python3
"""Exact all-pairs intersection cardinality on one MI355X.

Workload (the headline configuration that fits one GPU): dense contiguous
corpus, N = 32,768 rows x M = 1,048,576 bits (4 GiB packed u64), every bit
i.i.d. Bernoulli(0.05) from the splitmix64 stream Rng(2), generated on the
device; one step = every pair i < j, exact |Xi n Xj| written as a packed
upper-triangle u32 (536,854,528 counts, 2 GiB), inputs resident in HBM.

Prints ONE JSON line. `value` = exact pairs/s of the whole step.
`roofline` prices the dominant kernel (B x B) from HIP events recorded on the
engine's own stream around its launch. `cpu_baseline` times the reference's
own all-bitmap B x B kernel (compiled from its sources, oracle/_ref) on the
host cores over a bounded sample of the same rows.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "stormbitmaps-internal_amd")

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
MFMA_I8_PEAK_TOPS = 5000.0   # MI355X dense int8, TOP/s
MFMA_FP4_PEAK_TOPS = 10100.0  # MI355X dense fp4, TFLOP/s
HBM_PEAK_GBS = 8000.0


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


def cpu_baseline(sh, corpus, n_bits, target_s=12.0):
    """Reference all-bitmap B x B (kernels/cell_bb.cpp 'dense') with project
    threading over all pairs of the first R rows of the same corpus."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O  # CPU checker: the reference build lives under oracle/_ref
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
    r = int(min(8192, max(256, (2 * rate * target_s) ** 0.5)))
    s, t = run(r)
    pairs = r * (r - 1) // 2
    # The GPU's counts for the same pairs must agree.
    gpu = 0
    n = corpus.n
    for i in range(r - 1):
        off = i * n - i * (i + 1) // 2
        gpu += int(corpus.tri(off, r - i - 1).sum(dtype=np.uint64))
    return {"value": pairs / t, "unit": "pairs/s", "cores": threads, "kind": "reference",
            "sample": f"all {pairs} pairs of the first {r} rows of the same corpus, "
                      f"{t:.1f} s, reference bb 'dense' kernel, {threads} threads",
            "sample_sum": int(s), "sample_sum_matches_gpu": bool(gpu == s)}


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
    ctx = sh.Context(0)
    flags = 0 | 4 | 8   # bitmaps for dense rows, runs, EWAH
    t0 = time.perf_counter()
    corpus = ctx.from_lists(offs, pos, u, bitmap_policy=flags)
    t_up = time.perf_counter() - t0
    del pos
    pairs = n * (n - 1) // 2
    res = {}

    def timed(policy, **kw):
        corpus.allpairs(policy=policy, out="tri", tile=args.tile, **kw)   # warm-up
        ctx.sync()
        ts, st = [], None
        for _ in range(args.steps):
            t1 = time.perf_counter()
            st = corpus.allpairs(policy=policy, out="tri", tile=args.tile, **kw)
            ts.append(time.perf_counter() - t1)
        ms = 1000 * float(np.median(ts))
        return {"pairs_per_s": pairs / (ms / 1000), "ms": ms, "sum": st["sum"],
                "cell_tiles": st["cell_tiles"], "skipped": st["skipped"],
                "fallback_tiles": st["fallback_tiles"], "ms_select": st["ms_select"]}

    res["analytic_reference_rates"] = timed("analytic", use_graph=True)
    t0 = time.perf_counter()
    model = ctx.calibrate(corpus, tile=args.tile, max_tiles=args.sample_tiles, seed=7, repeats=2)
    t_cal = time.perf_counter() - t0
    res["analytic"] = timed("analytic", use_graph=True)
    t0 = time.perf_counter()
    X, Y, M, _ = ctx.selector_sample(corpus, tile=args.tile, max_tiles=args.sample_tiles,
                                     seed=11, repeats=2)
    t_sample = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = ctx.selector_train(X, Y, M, epochs=300, batch=128, lr=3e-3, seed=5)
    t_train = time.perf_counter() - t0
    res["learned"] = timed("learned", use_graph=True)
    if not args.no_allbitmap:
        res["all_bitmap"] = timed("all_bitmap")
    sums = {v["sum"] for v in res.values()}
    out = {
        "metric": "exact pairs/sec on 1 MI355X, learned vs analytic selection",
        "value": res["learned"]["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": res["learned"]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded host generator)",
        "config": {"workload": desc, "n_rows": n, "universe": u, "pairs": pairs, "tile": args.tile},
        "exact_sums_agree": len(sums) == 1, "sum": sums.pop() if len(sums) == 1 else None,
        "policies": res,
        "learned_over_analytic": res["learned"]["pairs_per_s"] / res["analytic"]["pairs_per_s"],
        "selector": {"sampled_tiles": int(len(X)), "train_loss": loss, "sample_s": t_sample,
                     "train_s": t_train},
        "calibrated_model": model, "calibrate_s": t_cal,
        "setup_s": {"generate": t_gen, "build_upload": t_up},
    }
    if "all_bitmap" in res:
        out["learned_over_all_bitmap"] = res["learned"]["pairs_per_s"] / res["all_bitmap"]["pairs_per_s"]
    corpus.free()
    ctx.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="dense2", choices=sorted(CONFIGS) + sorted(SPARSE_CONFIGS))
    ap.add_argument("--bb", default="mfma", choices=["mfma", "mfma_fp4", "valu"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rows", type=int, default=0, help="sparse configs: first N rows only")
    ap.add_argument("--tile", type=int, default=64)
    ap.add_argument("--sample-tiles", type=int, default=2048)
    ap.add_argument("--no-allbitmap", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(the all-pairs triangle would shard by tile rows, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    sh = load_storm_hip()
    if args.config in SPARSE_CONFIGS:
        sparse_bench(sh, args)
        return
    n, bits, p, seed, desc = CONFIGS[args.config]
    ctx = sh.Context(0)
    corpus = ctx.gen_bernoulli(n, bits, p, seed)
    for _ in range(args.warmup):
        corpus.allpairs(out="tri", bb=args.bb)
    ctx.sync()
    t0 = time.perf_counter()
    bb_ms, dev_ms, sums = [], [], set()
    for _ in range(args.steps):
        st = corpus.allpairs(out="tri", bb=args.bb)
        bb_ms.append(st["ms_bb"])
        dev_ms.append(st["ms_device"])
        sums.add(st["sum"])
    ctx.sync()
    wall = time.perf_counter() - t0
    ms_step = wall * 1000.0 / args.steps
    pairs = n * (n - 1) // 2
    kern_ms = float(np.mean(bb_ms))
    kernel = "bb_mfma" if args.bb == "mfma" else "bb_valu"
    if args.bb in ("mfma", "mfma_fp4"):
        fp4 = args.bb == "mfma_fp4"
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
        wp = pairs * ((bits + 63) // 64)
        valu_peak = 256 * 64 * 2.4e9 / 4            # 64-bit word-pairs/s (4 VALU ops each)
        achieved = wp / (kern_ms * 1e-3)
        roof = {"bound": "valu", "achieved": achieved / 1e12, "peak": valu_peak / 1e12,
                "unit": "T word-pairs/s", "frac": achieved / valu_peak,
                "traffic": traffic_from_profiles(kernel, args.config),
                "kernel": kernel, "kernel_ms": kern_ms,
                "algorithmic": "ceil(M/64) 64-bit AND+popcount per exact pair i<j"}
    out = {
        "metric": "exact pairs/sec on 1 MI355X",
        "value": pairs / (ms_step / 1000.0),
        "unit": "pairs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"mfma": "int8", "mfma_fp4": "fp4 (e2m1 0/1, f32 exact accumulate)"}.get(args.bb, "u32"),
        "data": "synthetic (device splitmix64 Bernoulli stream)",
        "config": {"workload": desc, "n_rows": n, "n_bits": bits, "density": p, "seed": seed,
                   "pairs": pairs, "bb_kernel": args.bb},
        "exact_sum": sums.pop() if len(sums) == 1 else None,
        "ms_device_per_step": float(np.mean(dev_ms)),
        "roofline": roof,
        "cpu_baseline": None if args.no_cpu else cpu_baseline(sh, corpus, bits),
    }
    corpus.free()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
