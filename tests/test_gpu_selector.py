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


def test_learned_policy_is_exact_and_beats_untrained_analytic(sh, ctx):
    u = 200_000
    n = 1600
    rows = _mixed(21, n, u)
    offs, pos = corpora.csr(rows)
    want_sum, want = O.allpairs_lists(offs, pos)
    # The untrained method: the analytic model with the reference's rates (it
    # also bounds the cells the learned selector may rank, at sampling and at
    # selection alike).
    ctx.set_cost_model([0.116, 0.30, 0.60, 0.55, 0.22, 0.35, 0.40, 0.30, 0.45, 0.50])
    with ctx.from_lists(offs, pos, u, bitmap_policy=0 | 4 | 8 | 16) as c:
        X, Y, M, idx = ctx.selector_sample(c, tile=64, max_tiles=400, seed=3, repeats=2)
        assert len(X) > 200
        assert M.any(axis=1).all()
        c.allpairs(policy="analytic", tile=64)
        T = (n + 63) // 64
        analytic = c.last_decisions(T * (T + 1) // 2)[idx]
        # Training and held-out tiles drawn from the same distribution (the
        # sample comes back in tile order, which runs from sparse to dense).
        order = np.random.default_rng(11).permutation(len(X))
        X, Y, M, idx, analytic = X[order], Y[order], M[order], idx[order], analytic[order]
        half = len(X) // 2
        loss = ctx.selector_train(X[:half], Y[:half], M[:half], epochs=400, batch=64, lr=3e-3,
                                  seed=7)
        print("train loss", loss, "labels per cell", M[:half].sum(axis=0), flush=True)
        # Held-out tiles: time of each method's choice against the
        # exhaustively timed best cell.
        pred = ctx.selector_predict(X[half:])
        Yh, Mh = Y[half:], M[half:].astype(bool)
        pred[~Mh] = np.inf
        rows = np.arange(len(Yh))
        t_best = np.exp2(np.where(Mh, Yh, np.inf).min(axis=1)).sum()
        t_learned = np.exp2(Yh[rows, pred.argmin(axis=1)]).sum()
        a = analytic[half:].astype(np.int64)
        assert Mh[rows, a].all()                  # analytic choices were timed
        t_analytic = np.exp2(Yh[rows, a]).sum()
        print("held-out regret: learned", t_learned / t_best - 1, "analytic",
              t_analytic / t_best - 1, flush=True)
        assert t_learned <= 1.05 * t_analytic
        st = c.allpairs(policy="learned", out="tri")
        assert st["sum"] == want_sum
        assert np.array_equal(c.tri(), want)
        st2 = c.allpairs(policy="analytic", out="tri")
        assert st2["sum"] == want_sum
