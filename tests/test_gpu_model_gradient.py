"""GPU: the model gradient (d logL / dP per branch and category, d logL / d pi
at the root) through the C ABI, and the GTR / restriction parameter
gradients built from it, against central finite differences of the CPU
oracle; invariants that tie it to the branch-length gradient; every
schedule and tile width; the batch form; the S > 8 error."""

import numpy as np
import pytest

from conftest import mcp, oracle, rel_err

pytestmark = pytest.mark.gpu
M = mcp.models


def _plan(w, schedule="levels"):
    p = mcp.Plan(w.ta.parent, w.x, len(w.rates), weights=w.weights)
    p.set_model(mcp.Model(w.U, w.D, w.Uinv, w.mu, w.rates, w.pi))
    p.set_schedule(schedule)
    return p


def _oracle_ll(w, pi, subst, blv):
    U, D, Uinv, mu = w.model_fn(pi, subst)
    return oracle.logl_grad(w.ta, w.x, U, D, Uinv, mu, w.rates, pi, blv, weights=w.weights,
                            grad=False)


def _fd(f, x, h_rel=1e-6):
    x = np.asarray(x, dtype=np.float64)
    g = np.zeros_like(x)
    for k in range(x.size):
        h = h_rel * max(1.0, abs(x[k]))
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        g[k] = (f(xp) - f(xm)) / (2 * h)
    return g


def _dPdt(w, blv):
    """dP/dt of every branch and category on the host: U diag(mu rho D e^a) Uinv."""
    out = np.zeros((len(blv), len(w.rates), w.S, w.S))
    for v, t in enumerate(blv):
        for r, rho in enumerate(w.rates):
            ex = w.mu * w.D * rho * np.exp(w.mu * t * w.D * rho)
            out[v, r] = (w.U * ex[None, :]) @ w.Uinv
    return out


@pytest.mark.parametrize("seed,n,C,R", [(701, 25, 600, 4), (702, 12, 300, 1)])
def test_gtr_parameter_gradient_vs_oracle_fd(seed, n, C, R):
    w = mcp.synthetic.make_workload("custom", seed=seed, n=n, S=4, C=C, R=R)
    blv = mcp.get_branchlength_vector(w.tree)
    p = _plan(w)
    print(f"gtr model gradient seed={seed} n={n} C={C} R={R}", flush=True)
    ll, g, dP, dpi = p.logl_grad_model(blv)
    ll2, g2 = p.logl_grad(blv)
    assert rel_err(ll, ll2) < 1e-14 and np.max(np.abs(g - g2)) <= 1e-12 * np.max(np.abs(g2))
    # sum_r <dlogL/dP_{v,r}, dP_{v,r}/dt> is the branch gradient
    gv = np.sum(dP * _dPdt(w, blv), axis=(1, 2, 3))
    assert np.max(np.abs(gv - g)) <= 1e-10 * np.max(np.abs(g)), np.max(np.abs(gv - g))
    # sum_s pi_s dlogL/dpi_s = R * sum_c w_c (each root term is log sum_s pi_s Ft_s)
    W = float(np.sum(w.weights)) if w.weights is not None else float(w.C)
    assert abs(float(w.pi @ dpi) - R * W) <= 1e-10 * R * W
    d_exch, d_pi = M.gtr_gradient(w.pi, w.subst_rates, w.U, w.D, w.Uinv, w.mu, blv, w.rates,
                                  dP, dpi)
    fd_exch = _fd(lambda e: _oracle_ll(w, w.pi, e, blv), w.subst_rates)
    fd_pi = _fd(lambda q: _oracle_ll(w, q, w.subst_rates, blv), w.pi)
    assert np.max(np.abs(d_exch - fd_exch)) <= 1e-6 * max(1.0, np.max(np.abs(fd_exch))), \
        (d_exch, fd_exch)
    assert np.max(np.abs(d_pi - fd_pi)) <= 1e-6 * max(1.0, np.max(np.abs(fd_pi))), (d_pi, fd_pi)


def test_restriction_parameter_gradient_vs_oracle_fd():
    w = mcp.synthetic.make_workload("custom", seed=711, n=20, S=2, C=500, R=4)
    blv = mcp.get_branchlength_vector(w.tree)
    p = _plan(w, "hybrid")
    print("restriction model gradient", flush=True)
    ll, g, dP, dpi = p.logl_grad_model(blv)
    d_pi = M.restriction_gradient(w.pi, w.U, w.D, w.Uinv, w.mu, blv, w.rates, dP, dpi)
    fd = _fd(lambda q: _oracle_ll(w, q, w.subst_rates, blv), w.pi)
    assert np.max(np.abs(d_pi - fd)) <= 1e-6 * max(1.0, np.max(np.abs(fd))), (d_pi, fd)


def test_model_gradient_same_for_every_schedule_and_tile_width():
    w = mcp.synthetic.make_workload("custom", seed=721, n=40, S=4, C=900, R=4)
    blv = mcp.get_branchlength_vector(w.tree)
    ref = None
    for sched in ("levels", "hybrid", "fused"):
        for cpt in (1, 2, 4):
            p = _plan(w, sched)
            p.set_tiles(4, cpt)
            print(f"model gradient schedule={sched} cpt={cpt}", flush=True)
            ll, g, dP, dpi = p.logl_grad_model(blv)
            if ref is None:
                ref = (ll, g, dP, dpi)
                continue
            assert rel_err(ll, ref[0]) < 1e-13
            assert np.max(np.abs(dP - ref[2])) <= 1e-11 * np.max(np.abs(ref[2]))
            assert np.max(np.abs(dpi - ref[3])) <= 1e-11 * np.max(np.abs(ref[3]))
            p.close()


def test_batch_model_gradient_matches_single_plans():
    ws = [mcp.synthetic.make_workload("custom", seed=731 + k, n=10 + 7 * k, S=S, C=200 + 50 * k,
                                      R=R)
          for k, (S, R) in enumerate([(4, 4), (2, 1), (4, 2)])]
    ds, items, blvs = [], [], []
    for k, w in enumerate(ws):
        ds.append({"x_leaves": w.x[:, :, w.ta.leaf_nums - 1], "leaf_nums": w.ta.leaf_nums})
        items.append({"dataset": k, "parent": w.ta.parent, "n_rates": len(w.rates)})
        blvs.append(mcp.get_branchlength_vector(w.tree))
    b = mcp.Batch(ds, items)
    for k, w in enumerate(ws):
        b.set_model(k, mcp.Model(w.U, w.D, w.Uinv, w.mu, w.rates, w.pi))
    print("batch model gradient", flush=True)
    ll, g, dPs, dpis = b.logl_grad_model(np.concatenate(blvs))
    for k, w in enumerate(ws):
        llk, gk, dPk, dpik = _plan(w).logl_grad_model(blvs[k])
        assert rel_err(ll[k], llk) < 1e-13
        assert np.max(np.abs(dPs[k] - dPk)) <= 1e-11 * np.max(np.abs(dPk))
        assert np.max(np.abs(dpis[k] - dpik)) <= 1e-11 * np.max(np.abs(dpik))


def test_model_gradient_rejects_twenty_states():
    w = mcp.synthetic.make_workload("custom", seed=741, n=6, S=20, C=64, R=1)
    p = _plan(w)
    with pytest.raises(mcp.McpError):
        p.logl_grad_model(mcp.get_branchlength_vector(w.tree))


@pytest.mark.parametrize("model", ["GTR", "JC"])
def test_phylodist_parameter_gradients_vs_oracle_fd(model):
    """gradlogpdf_params: branch lengths, category rates, base frequencies and
    exchangeabilities of a PhyloDist, each against oracle finite differences."""
    w = mcp.synthetic.make_workload("custom", seed=751, n=18, S=4, C=400, R=4)
    fn = getattr(mcp, model)
    d = mcp.PhyloDist(w.tree, w.pi, w.subst_rates, w.rates, fn)
    print(f"phylodist parameter gradients {model}", flush=True)
    ll, G = mcp.gradlogpdf_params(d, w.x)
    blv = mcp.get_branchlength_vector(w.tree)

    def ll_of(pi=w.pi, subst=w.subst_rates, rates=w.rates, b=blv):
        U, D, Uinv, mu = fn(pi, subst)
        return oracle.logl_grad(w.ta, w.x, U, D, Uinv, mu, rates, pi, b, grad=False)

    assert rel_err(ll, ll_of()) < 1e-12
    # the threaded oracle's sum changes in the last bits between calls, so a
    # central difference with h = 1e-6 carries noise of about 1e-15 |ll| / 1e-6
    noise = 1e-8 * abs(ll)
    for name, x0, f in [("base_freq", w.pi, lambda q: ll_of(pi=q)),
                        ("substitution_rates", w.subst_rates, lambda q: ll_of(subst=q)),
                        ("rates", w.rates, lambda q: ll_of(rates=q)),
                        ("branch_lengths", blv[:6], lambda q: ll_of(b=np.r_[q, blv[6:]]))]:
        fd = _fd(f, x0)
        got = G[name][:len(fd)]
        tol = 1e-6 * max(1.0, np.max(np.abs(fd))) + noise
        assert np.max(np.abs(got - fd)) <= tol, (name, got, fd)
