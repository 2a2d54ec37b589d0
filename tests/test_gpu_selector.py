"""The learned per-tile selector: device forward/backward/Adam against the
float64 CPU reference, training, and exactness of the learned policy."""
import ctypes as C

import numpy as np
import pytest

import corpora
import oracle_py as O

pytestmark = pytest.mark.gpu
P = O._p


def _shapes(sh):
    F, H1, H2, Ov, NP = sh.mlp_shape()
    return F, H1, H2, Ov, NP


def _split(p, F, H1, H2, Ov):
    sizes = [H1 * F, H1, H2 * H1, H2, Ov * H2, Ov]
    out, k = [], 0
    for s in sizes:
        out.append(np.ascontiguousarray(p[k:k + s], dtype=np.float64))
        k += s
    return out


def _orc_fwd_bwd(params, x, t, m, F, H1, H2, Ov):
    W = _split(params, F, H1, H2, Ov)
    G = [np.zeros_like(w) for w in W]
    y = np.zeros((len(x), Ov))
    L = O.oracle()
    x64 = np.ascontiguousarray(x, np.float64)
    t64 = np.ascontiguousarray(t, np.float64)
    loss = L.orc_mlp_forward_backward(F, H1, H2, Ov, len(x), P(x64), P(t64), P(m),
                                      *[P(w) for w in W], P(y), *[P(g) for g in G])
    return loss, np.concatenate(G), y


def test_device_gradients_match_float64_reference(sh, ctx):
    F, H1, H2, Ov, NP = _shapes(sh)
    rng = np.random.default_rng(11)
    params = rng.normal(0, 0.3, NP).astype(np.float32)
    norm = np.concatenate([np.zeros(F), np.ones(F), -1e30 * np.ones(F), 1e30 * np.ones(F)])
    ctx.selector_set_params(params, norm.astype(np.float32))
    for n in (1, 7, 300, 1000):
        x = rng.normal(size=(n, F)).astype(np.float32)
        t = rng.normal(size=(n, Ov)).astype(np.float32)
        m = (rng.random((n, Ov)) < 0.7).astype(np.uint8)
        g, loss = ctx.selector_grad(x, t, m)
        want_loss, want_g, _ = _orc_fwd_bwd(params.astype(np.float64), x, t, m, F, H1, H2, Ov)
        assert loss == pytest.approx(want_loss, rel=1e-4)
        scale = np.abs(want_g).max()
        assert np.abs(g - want_g).max() <= 1e-4 * scale + 1e-7, n


def test_device_adam_matches_published_update(sh, ctx):
    F, H1, H2, Ov, NP = _shapes(sh)
    rng = np.random.default_rng(12)
    params = rng.normal(0, 0.3, NP).astype(np.float32)
    norm = np.concatenate([np.zeros(F), np.ones(F), -1e30 * np.ones(F), 1e30 * np.ones(F)])
    ctx.selector_set_params(params, norm.astype(np.float32))   # resets m, v, step
    x = rng.normal(size=(64, F)).astype(np.float32)
    t = rng.normal(size=(64, Ov)).astype(np.float32)
    m = np.ones((64, Ov), np.uint8)
    p64 = params.astype(np.float64)
    mm = np.zeros(NP)
    vv = np.zeros(NP)
    for step in range(1, 4):
        g, _ = ctx.selector_grad(x, t, m)
        ctx.selector_adam_step(1e-2)
        g64 = g.astype(np.float64)
        O.oracle().orc_adam_step(NP, P(p64), P(g64), P(mm), P(vv), step, 1e-2, 0.9, 0.999, 1e-8)
        got, _ = ctx.selector_get_params()
        assert np.abs(got - p64).max() < 1e-5, step


def test_training_fits_and_predict_matches_reference_forward(sh, ctx):
    F, H1, H2, Ov, NP = _shapes(sh)
    rng = np.random.default_rng(13)
    n = 2000
    X = rng.normal(size=(n, F)).astype(np.float32) * 3 + 5
    w = rng.normal(size=(F, Ov))
    Y = (np.tanh((X - 5) / 3) @ w).astype(np.float32)
    M = np.ones((n, Ov), np.uint8)
    loss = ctx.selector_train(X, Y, M, epochs=150, batch=128, lr=3e-3, seed=5)
    base = 0.5 * float(np.mean((Y - Y.mean()) ** 2))
    assert loss < 0.2 * base, (loss, base)
    params, norm = ctx.selector_get_params()
    Xn = (X - norm[:F]) * norm[F:2 * F]
    _, _, y = _orc_fwd_bwd(params.astype(np.float64), Xn, np.zeros((n, Ov)),
                           np.zeros((n, Ov), np.uint8), F, H1, H2, Ov)
    got = ctx.selector_predict(X)
    assert np.abs(got - y).max() < 1e-3 * max(1.0, np.abs(y).max())


def _mixed(seed, n, u):
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        if i % 3 == 0:
            rows.append(corpora.runs_row(rng, u, int(rng.integers(50, 4000)), 200))
        else:
            k = int(np.exp(rng.random() * np.log(u * 0.2)))
            rows.append(corpora.uniform_row(rng, u, k))
    return rows


REF_RATES = [0.116, 0.30, 0.60, 0.55, 0.22, 0.35, 0.40, 0.30, 0.45, 0.50]


def _median_ms(c, policy, tile, n=5):
    import time
    c.allpairs(policy=policy, tile=tile, use_graph=True)
    ts = []
    for _ in range(n):
        t0 = time.perf_counter()
        c.allpairs(policy=policy, tile=tile, use_graph=True)
        ts.append(time.perf_counter() - t0)
    return 1000 * float(np.median(ts))


def test_explore_labels_every_tile_and_is_exact(sh, ctx):
    """An exploration call is an exact all-pairs call that labels every
    non-empty tile pair with the cell it ran and that cell's measured cost;
    other seeds draw other cells; batches label every member."""
    u, n = 200_000, 1600
    rows = _mixed(21, n, u)
    offs, pos = corpora.csr(rows)
    want_sum, _ = O.allpairs_lists(offs, pos)
    ctx.set_cost_model(REF_RATES)
    with ctx.from_lists(offs, pos, u, bitmap_policy=1 | 4 | 8 | 16 | 64) as c:
        for tile in (64, 256):
            T = (n + tile - 1) // tile
            c.allpairs(policy="analytic", tile=tile)
            dec = c.last_decisions(T * (T + 1) // 2)
            live = set(np.flatnonzero(dec != 10).tolist())
            seen = {}
            for seed in (1, 2, 3):
                X, Y, M, idx, s = ctx.selector_explore(c, tile=tile, seed=seed)
                assert s == want_sum, (tile, seed)
                assert (M.sum(axis=1) == 1).all()
                assert (Y[M > 0] > 0).all()
                assert len(set(idx.tolist())) == len(idx)
                assert set(idx.tolist()) == live, tile
                for t, cell in zip(idx.tolist(), M.argmax(axis=1).tolist()):
                    seen.setdefault(t, set()).add(cell)
            assert sum(len(v) > 1 for v in seen.values()) > len(seen) // 4
    # A batch labels the tiles of every member.
    rows2 = _mixed(22, 300, 50_000)
    o2, p2 = corpora.csr(rows2)
    s2, _ = O.allpairs_lists(o2, p2)
    with ctx.from_lists(offs, pos, u, bitmap_policy=0 | 4 | 8 | 16) as c1, \
            ctx.from_lists(o2, p2, 50_000, bitmap_policy=0 | 4 | 8 | 16) as c2, \
            ctx.batch([c1, c2]) as b:
        X, Y, M, idx, cidx, sums = b.explore(tile=256, seed=5)
        assert [int(x) for x in sums] == [want_sum, s2]
        assert set(cidx.tolist()) == {0, 1}
        assert (M.sum(axis=1) == 1).all()


def test_learned_policy_is_exact_and_beats_untrained_analytic(sh, ctx):
    """Run-heavy rows with every row's bitmap resident (cells B x B, B x R,
    R x R). The untrained method -- the analytic model with the reference's
    CPU rates -- against the learned selector trained on four exploration
    calls (costs measured in the production mix): learned must be strictly
    faster in production, by at least 20%, and both exact (total and every
    row sum against the column closed forms of the host generator's rows)."""
    n, u = 2048, 4_000_000
    offs, pos = sh.gen_corpus("runs", n, u, 20.0, 2000.0, 0.05, 71)
    total, rs = O.lists_closed_forms(offs, pos, u)
    ctx.set_cost_model(REF_RATES)
    with ctx.gen_device("runs", n, u, 20.0, 2000.0, 0.05, 71, bitmap_policy=1) as c:
        tile = 256
        parts = [ctx.selector_explore(c, tile=tile, seed=sd) for sd in (1, 2, 3, 4)]
        for p in parts:
            assert p[4] == total
        X, Y, M = (np.concatenate([p[k] for p in parts]) for k in range(3))
        loss = ctx.selector_train(X, Y, M, epochs=300, batch=128, lr=3e-3, seed=7)
        print("explored samples", len(X), "labels per cell", M.sum(axis=0), "loss", loss,
              flush=True)
        for policy in ("analytic", "learned"):
            st = c.allpairs(policy=policy, out="tri", tile=tile)
            assert st["sum"] == total, policy
            assert np.array_equal(c.row_sums(), rs), policy
            print(policy, st["cell_tiles"], flush=True)
        t_analytic = _median_ms(c, "analytic", tile)
        t_learned = _median_ms(c, "learned", tile)
        print(f"production: analytic (reference rates) {t_analytic:.2f} ms, learned "
              f"{t_learned:.2f} ms", flush=True)
        assert t_learned < 0.8 * t_analytic
