"""GPU: the device leapfrog (mcp_hmc_leapfrog) against its host
restatement with the CPU oracle's logL + gradient, graph replay, and NUTS on
the reference's simudata fixture."""

import numpy as np
import pytest

from conftest import load_fixture, mcp, oracle, rel_err

pytestmark = pytest.mark.gpu
S = mcp.sampler


def _simu_plan(schedule="levels"):
    t, x = load_fixture("simudata")
    ta = mcp.tree_arrays(t)
    U, D, Ui, mu = mcp.JC(np.ones(4) / 4)
    p = mcp.Plan(ta.parent, x, 1)
    p.set_model(mcp.Model(U, D, Ui, mu, [1.0], np.ones(4) / 4))
    p.set_schedule(schedule)
    return t, x, ta, (U, D, Ui, mu), p


def _oracle_logfgrad(ta, x, model):
    U, D, Ui, mu = model

    def f(blv):
        return oracle.logl_grad(ta, x, U, D, Ui, mu, [1.0], np.ones(4) / 4, blv)
    return f


@pytest.mark.parametrize("eps", [0.0, 0.01, 0.3])
def test_device_leapfrog_matches_host(eps):
    t, x, ta, model, p = _simu_plan()
    p.set_prior_exponential(0, 0.1)
    rng = np.random.default_rng(3)
    blv = mcp.get_branchlength_vector(t)
    r = rng.standard_normal(blv.shape[0]) * 2.0
    g = rng.standard_normal(blv.shape[0])
    delta = 0.02
    host = S.HostLeapfrog(_oracle_logfgrad(ta, x, model), prior_scale=0.1)
    ref = host.one(blv.copy(), r.copy(), g.copy(), eps, delta)
    xd, rd, gd = blv.copy(), r.copy(), g.copy()
    lp, ll = p.leapfrog(np.array([eps, delta]), xd, rd, gd)
    assert rel_err(xd, ref.x) < 1e-13
    assert np.array_equal(np.sign(rd), np.sign(ref.r))
    assert np.max(np.abs(rd - ref.r)) < 1e-8 * max(1.0, np.max(np.abs(ref.r)))
    assert np.max(np.abs(gd - ref.g)) < 1e-8 * np.max(np.abs(ref.g))
    assert rel_err(lp[0], ref.lf) < 1e-10


def test_leapfrog_graph_replay_matches_eager():
    t, x, ta, model, p = _simu_plan()
    p.set_prior_exponential(0, 0.1)
    rng = np.random.default_rng(4)
    blv = mcp.get_branchlength_vector(t)
    states = [(blv.copy(), rng.standard_normal(17), rng.standard_normal(17)) for _ in range(3)]
    eager = []
    for xs, rs, gs in states:
        a, b, c = xs.copy(), rs.copy(), gs.copy()
        lp, _ = p.leapfrog(np.array([0.05, 0.1]), a, b, c)
        eager.append((a, b, c, lp[0]))
    p.use_graph(True)
    for _ in range(2):
        for (xs, rs, gs), (ea, eb, ec, el) in zip(states, eager):
            a, b, c = xs.copy(), rs.copy(), gs.copy()
            lp, _ = p.leapfrog(np.array([0.05, 0.1]), a, b, c)
            assert np.array_equal(a, ea) and np.array_equal(b, eb) and np.array_equal(c, ec)
            assert lp[0] == el
    p.use_graph(False)


def test_nuts_on_simudata_runs_and_batches_exactly():
    t, x, ta, model, p = _simu_plan()
    p.set_prior_exponential(0, 0.1)
    p.use_graph(True)
    blv = mcp.get_branchlength_vector(t)
    lf = S.EngineLeapfrog(p, 1, [17])
    c = S.NUTSChain(blv, np.random.default_rng(11), target=0.8, tree_depth=6)
    draws, lfs, wall, rounds = S.run_nuts([c], lf, 40, n_adapt=20)
    d = draws[0]
    assert d.shape == (40, 17) and np.all(d > 0) and np.all(np.isfinite(lfs[0]))
    # two chains in one batch == each chain alone on a plan (deterministic kernels)
    leaf = ta.leaf_nums
    b = mcp.Batch([{"x_leaves": x[:, :, leaf - 1], "leaf_nums": leaf}],
                  [{"dataset": 0, "parent": ta.parent, "n_rates": 1}] * 2)
    U, D, Ui, mu = model
    for i in range(2):
        b.set_model(i, mcp.Model(U, D, Ui, mu, [1.0], np.ones(4) / 4))
        b.set_prior_exponential(i, 0.1)
    cs = [S.NUTSChain(blv, np.random.default_rng(s), tree_depth=5) for s in (21, 22)]
    both, *_ = S.run_nuts(cs, S.EngineLeapfrog(b, 2, [17, 17]), 15, n_adapt=5)
    for k, s in enumerate((21, 22)):
        c1 = S.NUTSChain(blv, np.random.default_rng(s), tree_depth=5)
        solo, *_ = S.run_nuts([c1], lf, 15, n_adapt=5)
        assert np.array_equal(solo[0], both[k])


@pytest.mark.parametrize("eps", [0.0, 0.01, 0.2])
def test_device_leapfrog_compound_dirichlet_prior_matches_host(eps):
    """Second half of the device leapfrog with the CompoundDirichlet prior
    (block sums of the tree length and of leaf / internal log lengths)
    against the host restatement with the oracle's prior."""
    t, x, ta, model, p = _simu_plan()
    args = (2.0, 1.5, 0.4, 0.7)
    p.set_prior_compound_dirichlet(0, *args)
    inl = mcp.priors.internal_external(t)

    def prior(b):
        return oracle.compound_dirichlet(b, inl, *args)

    rng = np.random.default_rng(9)
    blv = mcp.get_branchlength_vector(t)
    r = rng.standard_normal(blv.shape[0])
    g = rng.standard_normal(blv.shape[0])
    host = S.HostLeapfrog(_oracle_logfgrad(ta, x, model), prior=prior)
    ref = host.one(blv.copy(), r.copy(), g.copy(), eps, 0.02)
    xd, rd, gd = blv.copy(), r.copy(), g.copy()
    lp, _ = p.leapfrog(np.array([eps, 0.02]), xd, rd, gd)
    assert rel_err(xd, ref.x) < 1e-13
    assert np.max(np.abs(gd - ref.g)) < 1e-8 * np.max(np.abs(ref.g))
    assert np.max(np.abs(rd - ref.r)) < 1e-8 * max(1.0, np.max(np.abs(ref.r)))
    assert rel_err(lp[0], ref.lf) < 1e-10


def _simu_batch(nch, prior=0.1):
    t, x = load_fixture("simudata")
    ta = mcp.tree_arrays(t)
    U, D, Ui, mu = mcp.JC(np.ones(4) / 4)
    leaf = ta.leaf_nums
    b = mcp.Batch([{"x_leaves": x[:, :, leaf - 1], "leaf_nums": leaf}],
                  [{"dataset": 0, "parent": ta.parent, "n_rates": 1}] * nch)
    for i in range(nch):
        b.set_model(i, mcp.Model(U, D, Ui, mu, [1.0], np.ones(4) / 4))
        b.set_prior_exponential(i, prior)
    b.use_graph(True)
    return t, b


def test_native_nuts_on_engine_matches_host_driver_in_distribution():
    nch = 4
    t, b = _simu_batch(nch)
    blv = mcp.get_branchlength_vector(t)
    nuts = S.NativeNUTS([blv] * nch, [31, 32, 33, 34], target=0.8, tree_depth=6,
                        molifier=False)
    nuts.use_engine(b)
    nuts.run(50, adapt=True, keep_draws=False)
    dn, lpn, wall_n, rounds_n, _ = nuts.run(400)
    chains = [S.NUTSChain(blv, np.random.default_rng(40 + i), target=0.8, tree_depth=6,
                          molifier=False) for i in range(nch)]
    dh, _, wall_h, rounds_h = S.run_nuts(chains, S.EngineLeapfrog(b, nch, [17] * nch), 400,
                                         n_adapt=50)
    A, B = np.concatenate(dn), np.concatenate(dh)
    for k in range(17):
        se = np.hypot(S.mcse_bm(A[:, k]), S.mcse_bm(B[:, k]))
        assert abs(A[:, k].mean() - B[:, k].mean()) < 4.5 * se, k
    assert np.all(np.isfinite(lpn))
    print(f"native {wall_n / rounds_n * 1e3:.3f} ms/round, host {wall_h / rounds_h * 1e3:.3f}",
          flush=True)


def test_native_delayed_acceptance_on_simudata():
    nch, d, prior = 4, 17, 0.1
    t, b = _simu_batch(nch, prior)
    blv = mcp.get_branchlength_vector(t)
    rec = []
    lf = S.EngineLeapfrog(b, nch, [d] * nch, record=rec)
    warm = [S.NUTSChain(blv, np.random.default_rng(100 + i), target=0.8, tree_depth=6,
                        molifier=False) for i in range(nch)]
    S.run_nuts(warm, lf, 100, n_adapt=100)
    X = np.array([r[1] for r in rec])
    y = np.array([r[2] for r in rec])
    s = mcp.Surrogate(d, 256, 0, 1)
    s.fit_scaling(X, y)
    s.set_data(X, y)
    s.train(60, batch=256, lr=2e-3, lr_final=1e-5, seed=2)
    s.use_graph(True)

    def exact_logpost(Xs):
        return b.logl(Xs.reshape(-1)) + S.exponential_logprior(Xs, prior)
    nuts = S.NativeNUTS([c.x for c in warm], [51, 52, 53, 54], target=0.8, tree_depth=6,
                        molifier=False)
    nuts.set_state(eps=np.array([c.epsilon for c in warm]))
    nuts.use_surrogate(s, prior)
    draws, lps, wall, st = S.run_da_nuts_native(nuts, s, prior, exact_logpost, 800)
    assert st["accept_rate"] > 0.6, st
    nuts2 = S.NativeNUTS([c.x for c in warm], [61, 62, 63, 64], target=0.8, tree_depth=6,
                         molifier=False)
    nuts2.set_state(eps=np.array([c.epsilon for c in warm]))
    nuts2.use_engine(b)
    dp, *_ = nuts2.run(800)
    A, B = np.concatenate(draws), np.concatenate(dp)
    for k in range(d):
        se = np.hypot(S.mcse_bm(A[:, k]), S.mcse_bm(B[:, k]))
        assert abs(A[:, k].mean() - B[:, k].mean()) < 4.5 * se, k


def test_tree_length_posterior_of_the_reference_pnuts_test():
    """test/samplers/tree_samplers.jl:46-70: PNUTS on simudata under JC with a
    CompoundDirichlet(1, 1, 0.1, 1) tree prior, target 0.8, tree depth 5; the
    reference checks mean - sd <= 1.8 <= mean + sd for its tree-length summary.
    Here: the native driver on the engine with the prior in the device
    leapfrog (fixed topology and pi = 1/4: no NNI or pi moves)."""
    nch = 2
    t, x = load_fixture("simudata")
    ta = mcp.tree_arrays(t)
    U, D, Ui, mu = mcp.JC(np.ones(4) / 4)
    leaf = ta.leaf_nums
    b = mcp.Batch([{"x_leaves": x[:, :, leaf - 1], "leaf_nums": leaf}],
                  [{"dataset": 0, "parent": ta.parent, "n_rates": 1}] * nch)
    for i in range(nch):
        b.set_model(i, mcp.Model(U, D, Ui, mu, [1.0], np.ones(4) / 4))
        b.set_prior_compound_dirichlet(i, 1.0, 1.0, 0.1, 1.0)
    b.use_graph(True)
    blv = mcp.get_branchlength_vector(t)
    nuts = S.NativeNUTS([blv] * nch, [123, 124], target=0.8, tree_depth=5)
    nuts.use_engine(b)
    nuts.run(500, adapt=True, keep_draws=False)
    draws, lp, _, _, _ = nuts.run(1000)
    tl = np.concatenate([d.sum(axis=1) for d in draws])
    m, sd = tl.mean(), tl.std()
    print(f"tree length posterior {m:.3f} +- {sd:.3f}", flush=True)
    assert m - sd <= 1.8 <= m + sd, (m, sd)
    assert np.all(np.isfinite(lp))


def test_nni_refraction_on_the_engine_matches_the_oracle_and_samples():
    """refraction! with NNI attempts (refraction.jl:35-91) with the engine's
    logL: the same steps as with the CPU oracle (same RNG, so the same
    interchange choices and acceptances), and a short NUTS run that moves
    the topology and keeps every state consistent with its tree."""
    w = mcp.synthetic.make_workload("custom", seed=631, n=12, S=4, C=300, R=1)
    par0 = np.array(w.ta.parent, dtype=np.int32)
    N = len(par0)
    blv = mcp.get_branchlength_vector(w.tree).copy()
    internal = [v for v in range(1, N) if np.count_nonzero(par0 == v) == 2]
    for v in internal[:3]:
        blv[v - 1] = 2e-3                      # short internal branches reach zero
    plan = mcp.Plan(w.ta.parent, w.x, 1)
    plan.set_model(mcp.Model(w.U, w.D, w.Uinv, w.mu, w.rates, w.pi))

    class OraclePlan:
        def __init__(self):
            self.parent = par0.copy()

        def set_topology(self, parent):
            self.parent = np.array(parent, dtype=np.int32)

        def logl_grad(self, b):
            ta = mcp.tree_arrays(mcp.tree_from_parent(self.parent, b))
            return oracle.logl_grad(ta, w.x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, b)

        def logl(self, b):
            return self.logl_grad(b)[0]

    prior = lambda x: (float(np.sum(np.log(10.0) - 10.0 * x)), -10.0 * np.ones_like(x))
    rng = np.random.default_rng(5)
    x = S.TreePosition(blv, par0)
    lf0, g0 = plan.logl_grad(blv)
    dev = S.NNIRefractionLeapfrog(plan, prior=prior, rng=np.random.default_rng(9))
    host = S.NNIRefractionLeapfrog(OraclePlan(), prior=prior, rng=np.random.default_rng(9))
    sd, sh = S.HState(x, np.zeros(N - 1), g0, lf0), S.HState(x, np.zeros(N - 1), g0, lf0)
    for step in range(6):
        r = rng.standard_normal(N - 1) * 3.0
        print(f"nni leapfrog step {step}", flush=True)
        sd = dev.one(sd.x, r, sd.g, 0.01, 1e-300)
        sh = host.one(sh.x, r, sh.g, 0.01, 1e-300)
        assert np.array_equal(sd.x.parent, sh.x.parent)
        assert np.max(np.abs(np.asarray(sd.x) - np.asarray(sh.x))) <= 1e-10
        assert rel_err(sd.lf, sh.lf) < 1e-10
    assert dev.att_nni == host.att_nni and dev.nni == host.nni and dev.att_nni > 0

    chain = S.NUTSChain(S.TreePosition(blv, par0), np.random.default_rng(17), target=0.8,
                        tree_depth=5, molifier=False)
    lf_ = S.NNIRefractionLeapfrog(plan, prior=prior, rng=np.random.default_rng(19))
    draws, lfs, *_ = S.run_nuts([chain], lf_, 20, n_adapt=10)
    assert draws[0].shape == (20, N - 1) and np.all(draws[0] > 0)
    assert len(chain.topologies) == 20 and lf_.att_nni > 0
    plan.set_topology(chain.x.parent)
    ll = plan.logl(np.asarray(chain.x))
    assert rel_err(chain.lf, ll + prior(np.asarray(chain.x))[0]) < 1e-12
