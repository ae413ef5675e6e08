"""GPU: the surrogate network's BF16-MFMA kernels against the CPU oracle
(oracle/mlp_oracle.py) through the C ABI -- parameter gradients, inference,
input gradients, training, the surrogate leapfrog -- and delayed-acceptance
NUTS on the reference's simudata fixture."""

import numpy as np
import pytest

from conftest import load_fixture, mcp, mlp_oracle as mo

pytestmark = pytest.mark.gpu
S = mcp.sampler


def _random_surrogate(d, H, seed, scale=0.25):
    rng = np.random.default_rng(seed)
    s = mcp.Surrogate(d, H, 0, seed)
    flat = s.get_flat()
    b = s.unflatten(flat)
    b["W1"][:, :d] = rng.normal(0, scale, (H, d))
    b["b1"][:] = rng.normal(0, 0.2, H)
    b["W2"][:] = rng.normal(0, scale / 4, (H, H))
    b["b2"][:] = rng.normal(0, 0.2, H)
    b["w3"][:] = rng.normal(0, scale, H)
    b["b3"][0] = 0.3
    b["v"][:d] = rng.normal(0, 0.3, d)
    b["q"][:d] = rng.normal(0, 0.1, d)
    s.set_flat(flat)
    X = np.exp(rng.normal(-2.5, 0.7, (1, d)) + rng.normal(0, 0.3, (600, d)))
    L = np.log(X)
    s.set_scaling(L.mean(0), L.std(0), -1234.5, 6.5)
    return s, flat, X, rng


def _blocks_close(s, g, ref, rtol):
    gb, rb = s.unflatten(g.astype(np.float64)), s.unflatten(ref)
    for k in mo.BLOCKS:
        err = np.linalg.norm(gb[k] - rb[k])
        assert err <= rtol * np.linalg.norm(rb[k]) + 1e-30, (k, err, np.linalg.norm(rb[k]))


@pytest.mark.parametrize("d,H,n", [(17, 256, 100), (398, 256, 300), (70, 128, 64)])
def test_param_grad_matches_oracle(d, H, n):
    s, flat, X, rng = _random_surrogate(d, H, d + n)
    X = X[:n]
    logl = -1234.5 + 6.5 * rng.standard_normal(n)
    g, loss = s.param_grad(X, logl)
    args = (flat, H, s.D, X, logl, s.x_mean, s.x_sd, s.y_mean, s.y_sd)
    l_bf, g_bf = mo.param_grad(*args, mode="bf16")
    l_ex, g_ex = mo.param_grad(*args)
    # the device's arithmetic (BF16 operands, FP32 accumulation): agreement to
    # FP32 accumulation order
    _blocks_close(s, g, g_bf, 1e-3)
    assert abs(loss - l_bf) <= 1e-5 * l_bf
    # the exact FP64 network: agreement to BF16 operand precision
    _blocks_close(s, g, g_ex, 3e-2)
    assert abs(loss - l_ex) <= 1e-2 * l_ex
    # padding stays untouched by the gradient
    gb = s.unflatten(g)
    assert np.all(gb["W1"][:, d:] == 0) and np.all(gb["v"][d:] == 0) and np.all(gb["q"][d:] == 0)


@pytest.mark.parametrize("d,H,n", [(17, 256, 5), (398, 256, 130)])
def test_predict_and_input_gradient_match_oracle(d, H, n):
    s, flat, X, _ = _random_surrogate(d, H, 7 * d + n)
    X = X[:n]
    ll, g = s.predict(X, grad=True)
    args = (flat, H, s.D, X, s.x_mean, s.x_sd, s.y_mean, s.y_sd)
    ll_bf, g_bf = mo.predict(*args, mode="bf16")
    ll_ex, g_ex = mo.predict(*args)
    spread = s.y_sd * (1.0 + np.abs((ll_bf - s.y_mean) / s.y_sd))
    # FP32 tanhf / log on the device vs FP64 in the oracle put an operand on
    # the other side of a bf16 rounding tie about once in 3e4 elements: most
    # rows agree to FP32 accumulation order, every row to one flipped operand
    dev = np.abs(ll - ll_bf) / spread
    assert np.mean(dev <= 1e-5) >= 0.95 and np.all(dev <= 1e-3), np.sort(dev)[-5:]
    assert np.linalg.norm(g - g_bf) <= 1e-3 * np.linalg.norm(g_bf)
    # against the exact FP64 network: BF16 operand precision (unit roundoff
    # 2^-9) over K = d and H terms, in standardised units
    ey = (ll - ll_ex) / s.y_sd
    assert np.sqrt(np.mean(ey ** 2)) <= 0.03 and np.max(np.abs(ey)) <= 0.08
    assert np.linalg.norm(g - g_ex) <= 3e-2 * np.linalg.norm(g_ex)
    # predict without the gradient gives the same values
    assert np.array_equal(s.predict(X), ll)


def test_training_fits_a_smooth_target_and_is_deterministic():
    d, H, n = 24, 256, 6000
    rng = np.random.default_rng(5)
    A = rng.normal(0, 1, (d, d)) / np.sqrt(d)
    X = np.exp(rng.normal(-2.0, 0.5, (n, d)))

    def target(X):
        z = (np.log(X) + 2.0) / 0.5
        u = z @ A
        return -300.0 - 0.5 * np.sum(u * u, axis=1) + 2.0 * np.tanh(z[:, 0] * z[:, 1])

    y = target(X)
    Xte = np.exp(rng.normal(-2.0, 0.5, (1000, d)))
    yte = target(Xte)
    runs = []
    for _ in range(2):
        s = mcp.Surrogate(d, H, 0, 11)
        s.fit_scaling(X, y)
        s.set_data(X, y)
        losses = s.train(200, batch=256, lr=2e-3, lr_final=1e-5, seed=3)
        runs.append((s.get_flat(), losses, s.predict(Xte)))
    (p1, l1, pr1), (p2, l2, pr2) = runs
    assert np.array_equal(p1, p2) and np.array_equal(l1, l2)   # deterministic kernels
    assert l1[-1] < 0.01 * l1[0], (l1[0], l1[-1])
    r2 = 1.0 - np.mean((pr1 - yte) ** 2) / np.var(yte)
    assert r2 > 0.95, r2


@pytest.mark.parametrize("lam", [0.0, 3.0])
def test_training_graph_replay_matches_eager_steps(lam):
    # n = 3 full minibatches + a remainder: the graph serves the full ones
    # (offset and Adam rate from the device step counter), the remainder runs
    # eagerly; turning the graph off must give the same bits.
    d, H, n, batch = 21, 128, 3 * 128 + 37, 128
    rng = np.random.default_rng(9)
    X = np.exp(rng.normal(-2.0, 0.5, (n, d)))
    y = -200.0 - np.sum(np.log(X) ** 2, axis=1)
    G = -2.0 * np.log(X) / X
    runs = []
    for graph in (True, False):
        s = mcp.Surrogate(d, H, 0, 4)
        s.train_graph(graph)
        s.fit_scaling(X, y)
        s.set_data(X, y, G)
        l1 = s.train(3, batch=batch, lr=2e-3, lr_final=1e-4, seed=8, grad_weight=lam)
        # a second call continues the Adam step count and reuses the graph
        l2 = s.train(2, batch=batch, lr=1e-3, lr_final=1e-4, seed=9, grad_weight=lam)
        runs.append((s.get_flat(), l1, l2))
    (p1, a1, b1), (p2, a2, b2) = runs
    assert np.all(np.isfinite(p1)) and a1[-1] < a1[0]
    assert np.array_equal(a1, a2) and np.array_equal(b1, b2)
    assert np.array_equal(p1, p2)


def test_surrogate_leapfrog_matches_host_restatement():
    d, H = 17, 256
    s, flat, X, rng = _random_surrogate(d, H, 99)

    def logfgrad(x):
        ll, g = s.predict(x, grad=True)
        return float(ll[0]), g[0]

    host = S.HostLeapfrog(logfgrad, prior_scale=0.1)
    chains = 3
    x = np.ascontiguousarray(X[:chains])
    r = rng.standard_normal((chains, d)) * 2.0
    g = rng.standard_normal((chains, d))
    ed = np.array([0.0, 0.02, 0.01, 0.02, 0.3, 0.05])
    refs = [host.one(x[i].copy(), r[i].copy(), g[i].copy(), ed[2 * i], ed[2 * i + 1])
            for i in range(chains)]
    xd, rd, gd = x.copy(), r.copy(), g.copy()
    lp = s.leapfrog(ed, xd, rd, gd, prior_scale=0.1)
    for i, ref in enumerate(refs):
        assert np.max(np.abs(xd[i] - ref.x) / ref.x) < 1e-13
        assert np.max(np.abs(gd[i] - ref.g)) <= 1e-9 * np.max(np.abs(ref.g))
        assert np.max(np.abs(rd[i] - ref.r)) <= 1e-9 * max(1.0, np.max(np.abs(ref.r)))
        assert abs(lp[i] - ref.lf) <= 1e-9 * abs(ref.lf)


def test_surrogate_leapfrog_graph_replay_matches_eager():
    d, H = 40, 256
    s, flat, X, rng = _random_surrogate(d, H, 5)
    chains = 5
    states = [(np.ascontiguousarray(X[5 * k:5 * k + chains]), rng.standard_normal((chains, d)),
               rng.standard_normal((chains, d))) for k in range(3)]
    ed = np.tile([0.02, 0.04], chains)
    eager = []
    for xs, rs, gs in states:
        a, b, c = xs.copy(), rs.copy(), gs.copy()
        eager.append((a, b, c, s.leapfrog(ed, a, b, c, 0.1)))
    s.use_graph(True)
    for _ in range(2):
        for (xs, rs, gs), (ea, eb, ec, el) in zip(states, eager):
            a, b, c = xs.copy(), rs.copy(), gs.copy()
            lp = s.leapfrog(ed, a, b, c, 0.1)
            assert np.array_equal(a, ea) and np.array_equal(b, eb) and np.array_equal(c, ec)
            assert np.array_equal(lp, el)
    s.use_graph(False)


def test_surrogate_argument_errors():
    with pytest.raises(mcp.McpArgumentError):
        mcp.Surrogate(10, 100)                       # hidden not a multiple of 64
    s = mcp.Surrogate(5, 64)
    with pytest.raises(mcp.McpError) as e:
        s.train(1)                                   # no data
    assert e.value.code == mcp._lib.MCP_ERR_STATE
    with pytest.raises(mcp.McpArgumentError):
        s.predict(np.array([[0.1, 0.2, -0.1, 0.3, 0.1]]))   # non-positive branch length
    with pytest.raises(mcp.McpArgumentError):
        s.set_scaling(np.zeros(5), np.zeros(5), 0.0, 1.0)   # zero feature sd
    with pytest.raises(ValueError):
        s.predict(np.ones((2, 4)))


def _simu_setup(n_chains):
    t, x = load_fixture("simudata")
    ta = mcp.tree_arrays(t)
    U, D, Ui, mu = mcp.JC(np.ones(4) / 4)
    leaf = ta.leaf_nums
    b = mcp.Batch([{"x_leaves": x[:, :, leaf - 1], "leaf_nums": leaf}],
                  [{"dataset": 0, "parent": ta.parent, "n_rates": 1}] * n_chains)
    for i in range(n_chains):
        b.set_model(i, mcp.Model(U, D, Ui, mu, [1.0], np.ones(4) / 4))
        b.set_prior_exponential(i, 0.1)
    b.use_graph(True)
    return t, b


def test_delayed_acceptance_matches_plain_nuts_and_saves_exact_evaluations():
    """Surrogate trained on the sampler's own (state, logL) pairs; DA-NUTS
    reproduces plain NUTS's posterior means within MCSE while evaluating the
    exact likelihood once per iteration, and the trained network accepts far
    more second-stage proposals than the untrained one."""
    nch, d, prior = 4, 17, 0.1
    t, b = _simu_setup(nch)
    blv = mcp.get_branchlength_vector(t)
    rec = []
    lf = S.EngineLeapfrog(b, nch, [d] * nch, record=rec)
    # molifier off: with the reference's molifier the zero-length step at the
    # start of each transition moves branches shorter than 2 eps, the
    # transition is not reversible there and the two samplers (different eps)
    # would target slightly different distributions
    warm = [S.NUTSChain(blv, np.random.default_rng(100 + i), target=0.8, tree_depth=6,
                        molifier=False) for i in range(nch)]
    S.run_nuts(warm, lf, 150, n_adapt=100)
    X = np.array([r[1] for r in rec])
    y = np.array([r[2] for r in rec])
    assert X.shape[0] > 5000

    def exact_logpost(Xs):
        return b.logl(Xs.reshape(-1)) + S.exponential_logprior(Xs, prior)

    # plain NUTS reference run (exact leapfrogs)
    lf.record = None
    plain = [S.NUTSChain(w.x, np.random.default_rng(200 + i), target=0.8, tree_depth=6,
                         epsilon=w.epsilon, molifier=False) for i, w in enumerate(warm)]
    dp, _, _, _ = S.run_nuts(plain, lf, 1000, n_adapt=50)
    exact_plain = sum(c.leapfrogs for c in plain)

    results = {}
    for trained in (True, False):
        s = mcp.Surrogate(d, 256, 0, 1)
        s.fit_scaling(X, y)
        if trained:
            s.set_data(X, y)
            s.train(60, batch=256, lr=2e-3, lr_final=1e-5, seed=2)
        s.use_graph(True)
        chains = [S.NUTSChain(w.x, np.random.default_rng(300 + i), target=0.8, tree_depth=6,
                              epsilon=w.epsilon, molifier=False) for i, w in enumerate(warm)]
        n_iter = 1000 if trained else 200
        draws, lps, wall, st = S.run_da_nuts(chains, S.SurrogateLeapfrog(s, nch, prior),
                                             exact_logpost, n_iter, n_adapt=50)
        results[trained] = (draws, st)
    draws, st = results[True]
    # The second-stage acceptance rate is a property of one run: it moves by
    # a few points when the chains' trajectories change (any change in the
    # rounding of the exact gradients changes them; 0.69 and above 0.7 have
    # both been seen on this workload).  0.6 still separates a trained
    # surrogate from an untrained one, which the next line checks directly.
    assert st["accept_rate"] > 0.6, st["accept_rate"]
    assert results[False][1]["accept_rate"] < 0.5 * st["accept_rate"]
    # posterior means agree within MCSE (pooled chains, batch means)
    A = np.concatenate(draws)
    B = np.concatenate(dp)
    for k in range(d):
        se = np.hypot(S.mcse_bm(A[:, k]), S.mcse_bm(B[:, k]))
        assert abs(A[:, k].mean() - B[:, k].mean()) <= 4.5 * se, (k, A[:, k].mean(),
                                                                  B[:, k].mean(), se)
    # exact evaluations per effective sample, with the median and with the
    # minimum ESS over branches.  On this 17-branch tree an exact leapfrog is
    # cheap and the gain is 2-3.5x across runs (each ratio of two batch-means
    # ESS estimates from one run moves by tens of percent); the bar is a clear
    # saving.  The 200-taxon bench workload shows the ESS-per-second gain.
    ess_da = np.array([S.ess_bm(np.concatenate([c[:, k] for c in draws])) for k in range(d)])
    ess_pl = np.array([S.ess_bm(np.concatenate([c[:, k] for c in dp])) for k in range(d)])
    gain_med = (exact_plain / np.median(ess_pl)) / (st["exact_evals"] / np.median(ess_da))
    gain_min = (exact_plain / ess_pl.min()) / (st["exact_evals"] / ess_da.min())
    print(f"DA gain in exact evaluations per effective sample: median {gain_med:.2f}, "
          f"min {gain_min:.2f}; accept {st['accept_rate']:.3f}", flush=True)
    assert gain_med > 1.5 and gain_min > 1.5, (gain_med, gain_min)


@pytest.mark.parametrize("d,H,n,lam", [(17, 256, 100, 2.0), (398, 256, 200, 50.0)])
def test_gradient_matching_param_grad_matches_oracle(d, H, n, lam):
    s, flat, X, rng = _random_surrogate(d, H, 3 * d + n)
    X = X[:n]
    logl = -1234.5 + 6.5 * rng.standard_normal(n)
    dlogl = rng.normal(0, 30, (n, d))
    g, loss = s.param_grad(X, logl, grad=dlogl, grad_weight=lam)
    args = (flat, H, s.D, X, logl, s.x_mean, s.x_sd, s.y_mean, s.y_sd)
    l_bf, g_bf = mo.param_grad(*args, mode="bf16", dlogl=dlogl, grad_weight=lam)
    l_ex, g_ex = mo.param_grad(*args, dlogl=dlogl, grad_weight=lam)
    _blocks_close(s, g, g_bf, 2e-3)
    assert abs(loss - l_bf) <= 1e-4 * l_bf
    _blocks_close(s, g, g_ex, 5e-2)
    # weight 0 with gradients given equals the value-only loss
    g0, l0 = s.param_grad(X, logl, grad=dlogl, grad_weight=0.0)
    g1, l1 = s.param_grad(X, logl)
    assert np.array_equal(g0, g1) and l0 == l1
