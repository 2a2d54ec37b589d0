"""The benchmark workloads at scale: size-independent checks.

Rows of the skewed (Zipf) and run-heavy workloads are generated from
independent per-row streams, so the first N rows of a workload are a fixed
corpus. At these sizes the full per-pair oracle is out of reach on the CPU;
the checks are exact nonetheless:
  * every row sum sum_j |Xi n Xj| equals sum_{b in Xi} (c_b - 1) from the
    column counts c_b (closed form),
  * the total equals sum_b C(c_b, 2),
  * a seeded sample of pairs equals the oracle's merge count,
  * every policy returns the same per-pair counts.
"""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu

WORKLOADS = {
    "zipf3": ("zipf", 3866847, 1.1, 386684, 0.0, 3, 8192),
    "runs4": ("runs", 36974578, 20.0, 2000.0, 0.05, 4, 2048),
}


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_workload_prefix_exact(sh, ctx, name):
    kind, u, p1, p2, p3, seed, n = WORKLOADS[name]
    print("case", name, n, flush=True)
    offs, pos = sh.gen_corpus(kind, n, u, p1, p2, p3, seed)
    total, row_sums = O.lists_closed_forms(offs, pos, u)
    rng = np.random.default_rng(99)
    ii = rng.integers(0, n, 2000)
    jj = rng.integers(0, n, 2000)
    L = O.oracle()
    with ctx.from_lists(offs, pos, u, bitmap_policy=0 | 4 | 8) as c:
        ref_tri = None
        for policy, kw in [("analytic", {}), ("fixed", {"fixed_cell": 4}),
                           ("analytic", {"use_graph": True})]:
            st = c.allpairs(policy=policy, out="tri", **kw)
            assert st["sum"] == total, (policy, kw)
            assert np.array_equal(c.row_sums(), row_sums), (policy, kw)
            tri = c.tri()
            if ref_tri is None:
                ref_tri = tri
                for i, j in zip(ii, jj):
                    if i == j:
                        continue
                    a, b = (i, j) if i < j else (j, i)
                    A = pos[offs[a]:offs[a + 1]]
                    B = pos[offs[b]:offs[b + 1]]
                    want = L.orc_intersect_merge(O._p(A), len(A), O._p(B), len(B))
                    assert tri[a * n - a * (a + 1) // 2 + (b - a - 1)] == want
            else:
                assert np.array_equal(tri, ref_tri), (policy, kw)


def test_dense_config_full_size(ctx):
    """The dense contiguous workload at its stated size: N = 32,768 rows x
    M = 1,048,576 bits, Bernoulli(0.05), seed 2, generated on the device.
    Under the fp4 and the int8 matrix-core kernels: the total equals
    sum_b C(c_b, 2) and every row sum equals sum_{b in Xi} (c_b - 1), from the
    column counts of the downloaded rows (the generator itself is checked
    against the oracle stream on a band of rows); three bands of the triangle
    (first, middle and last rows, every j > i) equal the oracle pair by pair."""
    n, bits, p, seed = 32768, 1 << 20, 0.05, 2
    nw = bits // 64
    print("case dense config full size", flush=True)
    with ctx.gen_bernoulli(n, bits, p, seed) as c:
        words = c.download_words(0, n, nw)
        assert np.array_equal(words[1000:1004], O.bernoulli_rows(seed, p, bits, 1000, 4))
        total, row_sums = O.column_closed_forms(words, threads=16)
        bands = [(0, 32), (n // 2 - 16, n // 2 + 16), (n - 32, n)]
        want_bands = [O.band_words(words, r0, r1, threads=16)[1] for r0, r1 in bands]
        del words
        for bb in ("auto", "mfma"):
            print("case dense config full size", bb, flush=True)
            st = c.allpairs(out="tri", bb=bb)
            assert st["pairs"] == n * (n - 1) // 2
            assert st["sum"] == total, bb
            assert np.array_equal(c.row_sums(), row_sums), bb
            for (r0, r1), want in zip(bands, want_bands):
                lo = r0 * n - r0 * (r0 + 1) // 2
                assert np.array_equal(c.tri(lo, len(want)), want), (bb, r0, r1)


def _instance_rows(c, inst):
    """Rows of a generated instance as CSR from its stored representation."""
    a = c.arenas()
    n, u = inst["n"], inst["universe"]
    if len(a["list"]) or inst["kind"] == "bernoulli" and inst["density"] * 32 < 1:
        return a["list_off"].astype(np.uint64), a["list"]
    if inst["kind"] == "runs":
        rows = []
        for r in range(n):
            lo, hi = int(a["run_off"][r]), int(a["run_off"][r + 1])
            rows.append(np.concatenate([np.arange(s, e, dtype=np.uint32) for s, e in
                                        zip(a["run_start"][lo:hi], a["run_end"][lo:hi])])
                        if hi > lo else np.zeros(0, np.uint32))
    else:
        nw = (u + 63) // 64
        ld = len(a["bitmap"]) // n
        bm = a["bitmap"].reshape(n, ld)[a["bm_row"].astype(np.int64), :nw]
        rows = [np.flatnonzero(np.unpackbits(w.view(np.uint8), bitorder="little")[:u])
                .astype(np.uint32) for w in bm]
    offs = np.zeros(n + 1, np.uint64)
    offs[1:] = np.cumsum([len(r) for r in rows])
    return offs, (np.concatenate(rows).astype(np.uint32) if offs[-1] else np.zeros(0, np.uint32))


def test_batched_instances_subset(sh, ctx):
    """The batched workload's own instances (seeded table of 512): the twelve
    smallest and every 24th of the rest up to 2 GiB each, generated on the
    device and run as one batch. Every instance's sum equals its column closed
    form (computed on the device without any cell, and on the CPU for the
    small ones), every row sum equals sum_{b in Xi} (c_b - 1), and the small
    instances equal the oracle pair by pair and the host generator row by
    row."""
    inst = sh.batched_instances(512, 5)
    by_size = sorted(inst, key=lambda x: x["est_bytes"])
    chosen = {x["k"]: x for x in by_size[:12]}
    for x in inst[::24]:
        if x["est_bytes"] <= 2 << 30:
            chosen[x["k"]] = x
    chosen = [chosen[k] for k in sorted(chosen)]
    corpora, wants = [], []
    for x in chosen:
        print("case batched instance", x["k"], x["kind"], x["n"], x["universe"],
              f"{x['density']:.3g}", flush=True)
        c = sh.gen_instance(ctx, x)
        total, rs = c.closed_form()
        corpora.append(c)
        wants.append((total, rs))
        small = x["n"] * x["universe"] * x["density"] <= 3e6 and x["n"] <= 1500
        if small:
            offs, pos = _instance_rows(c, x)
            t2, rs2 = O.lists_closed_forms(offs, pos, x["universe"])
            assert t2 == total and np.array_equal(rs2, rs), x["k"]
            kind, n, u, p1, p2, p3, seed, _ = x["gen"]
            if kind == "runs" or p1 * 32 < 1:
                ho, hp = sh.gen_corpus(kind, n, u, p1, p2, p3, seed)
                assert np.array_equal(ho, offs) and np.array_equal(hp, pos), x["k"]
            else:
                nw = (u + 63) // 64
                words = O.bernoulli_rows(seed, p1, u, 0, n)
                ld_rows = [np.flatnonzero(np.unpackbits(w.view(np.uint8),
                                                        bitorder="little")[:u]) for w in words]
                assert all(np.array_equal(a, pos[int(offs[r]):int(offs[r + 1])])
                           for r, a in enumerate(ld_rows)), x["k"]
            x["oracle"] = O.allpairs_lists(offs, pos)
    print("case batched run", len(corpora), flush=True)
    with ctx.batch(corpora) as b:
        for policy in ("analytic", "fixed"):
            st, sums, lat = b.allpairs(policy=policy, out="tri", tile=256, fixed_cell=7,
                                       use_graph=True)
            for c, x, (total, rs), s in zip(corpora, chosen, wants, sums):
                assert int(s) == total, (policy, x["k"])
                assert np.array_equal(c.row_sums(), rs), (policy, x["k"])
                if "oracle" in x:
                    assert np.array_equal(c.tri(), x["oracle"][1]), (policy, x["k"])
            assert (lat > 0).sum() == sum(1 for x in chosen if x["n"] >= 2)
    for c in corpora:
        c.free()


def test_batched_instances_all_512(sh, ctx):
    """The batched workload in full: all 512 instances, generated on the device
    in their stored representations, run in batches of about 150 GiB
    (estimated); every instance's sum equals its column closed form (computed
    on the device from the rows, no cell). The analytic model prices only
    B x B, S x S and R x R here (the cells the device-calibrated model picks
    on this workload), so the whole run takes about a minute."""
    inst = sh.batched_instances(512, 5)
    budget = 150 << 30
    batches, cur, used = [], [], 0
    for x in inst:
        if cur and used + x["est_bytes"] > budget:
            batches.append(cur)
            cur, used = [], 0
        cur.append(x)
        used += x["est_bytes"]
    if cur:
        batches.append(cur)
    saved = ctx.get_cost_model()
    ctx.set_cost_model([0.116, 0, 0, 0, 0.22, 0, 0, 0.30, 0, 0])
    checked = 0
    try:
        for bi, group in enumerate(batches):
            print("case batched instances batch", bi, "of", len(batches), ":", len(group),
                  "instances", flush=True)
            corpora = [sh.gen_instance(ctx, x) for x in group]
            wants = [int(c.closed_form(row_sums=False)[0]) for c in corpora]
            with ctx.batch(corpora) as b:
                st, sums, lat = b.allpairs(policy="analytic", out="sum", tile=256)
                for x, s_, w in zip(group, sums, wants):
                    assert int(s_) == w, (x["k"], int(s_), w)
                    checked += 1
            for c in corpora:
                c.free()
            print("  done, cells", st["cell_tiles"], flush=True)
    finally:
        ctx.put_cost_model(saved)
    assert checked == 512
