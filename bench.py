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
MFMA_I8_PEAK_TOPS = 5000.0   # MI355X dense int8, TOP/s
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="dense2", choices=sorted(CONFIGS))
    ap.add_argument("--bb", default="mfma", choices=["mfma", "valu"])
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(the all-pairs triangle would shard by tile rows, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    sh = load_storm_hip()
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
    if args.bb == "mfma":
        ops = 2.0 * pairs * bits                      # int8 MACs x 2, exact pairs only
        achieved = ops / (kern_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": MFMA_I8_PEAK_TOPS,
                "unit": "TFLOP/s", "frac": achieved / MFMA_I8_PEAK_TOPS,
                "traffic": traffic_from_profiles(kernel, args.config),
                "kernel": kernel, "kernel_ms": kern_ms,
                "algorithmic": "2*M int8 ops per exact pair i<j (0/1 operands, i32 accumulate)"}
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
        "dtype": "int8" if args.bb == "mfma" else "u32",
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
