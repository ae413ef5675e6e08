"""Pins the CPU oracle (oracle/felsenstein_oracle.c) before anything is
compared against it.

Evidence used, each chosen so that a plausible mistake fails at least one:
  * the reference's golden values (test/likelihood/felsenstein.jl:20,35,37):
    primates logL, simudata gradient (fixes node numbering, the adjoint, the
    sibling product, P vs P^T);
  * brute-force enumeration of internal states on small trees with
    non-symmetric P (GTR, unequal pi), gaps, R > 1 (fixes the forward pass,
    the per-(site, category) log semantics, transposition, scaling);
  * brute-force derivatives (fixes the gradient exactly, not just to FD
    precision) and central finite differences (h = 1e-6);
  * the pulley principle at a binary root (reversible models);
  * pattern weights == duplicated columns; R identical categories == R x ll.
"""

import numpy as np
import pytest

from conftest import (PRIMATES_LL, SIMU_GRAD, SIMU_LL, load_fixture, mcp, oracle, rel_err)
from _brute import brute_logl, rate_matrix


def _jc4():
    return mcp.JC(np.ones(4) / 4)


def _oracle_fixture(name):
    t, x = load_fixture(name)
    U, D, Ui, mu = _jc4()
    ta = mcp.tree_arrays(t)
    return ta, x, oracle.logl_grad(ta, x, U, D, Ui, mu, [1.0], np.ones(4) / 4,
                                   mcp.get_branchlength_vector(t))


def test_primates_golden():
    _, _, (ll, _) = _oracle_fixture("primates")
    assert rel_err(ll, PRIMATES_LL) < 1e-12


def test_simudata_gradient_golden():
    # numbering: leaves 1..n in Newick order, internal nodes n+1..N post-order
    _, _, (ll, g) = _oracle_fixture("simudata")
    assert g.shape == (17,)
    assert rel_err(g, SIMU_GRAD) < 1e-11


def test_simudata_loglik_brute_force():
    """The reference's simudata logL golden (-738.7363911756175) agrees with
    the exact value only to 2e-9 relative (its test uses isapprox, rtol
    1.5e-8).  Brute-force enumeration fixes the exact value, and the oracle
    matches that to 1e-12."""
    ta, x, (ll, _) = _oracle_fixture("simudata")
    Q, mu = rate_matrix("JC", np.ones(4) / 4, None)
    t, _ = load_fixture("simudata")
    exact = brute_logl(ta, x, Q, mu, [1.0], np.ones(4) / 4, mcp.get_branchlength_vector(t))
    assert rel_err(ll, exact) < 1e-12
    assert rel_err(ll, SIMU_LL) < 1.5e-8


def _small_case(seed, S, R, n, gaps=True, topology="coalescent"):
    rng = np.random.default_rng(seed)
    w = mcp.synthetic.make_workload("custom", seed=seed, n=n, S=S, C=37, R=R, topology=topology)
    x = np.array(w.x, order="F")
    if gaps:
        leaves = w.ta.leaf_nums
        for _ in range(8):
            v = int(rng.choice(leaves)) - 1
            x[:, int(rng.integers(x.shape[1])), v] = 1.0
    kind = "Restriction" if S == 2 else "GTR"
    return w, x, kind


@pytest.mark.parametrize("seed,S,R,n,topo", [
    (1, 2, 1, 5, "coalescent"), (2, 4, 1, 5, "yule"), (3, 4, 3, 4, "caterpillar"),
    (4, 3, 2, 5, "coalescent"), (5, 2, 4, 6, "caterpillar"),
])
def test_brute_force_small_trees(seed, S, R, n, topo):
    w, x, kind = _small_case(seed, S, R, n, topology=topo)
    blv = mcp.get_branchlength_vector(w.tree)
    ll, g = oracle.logl_grad(w.ta, x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, blv)
    Q, mu = rate_matrix(kind, w.pi, w.subst_rates)
    llb, gb = brute_logl(w.ta, x, Q, mu, w.rates, w.pi, blv, grad=True)
    assert rel_err(ll, llb) < 1e-11
    assert np.max(np.abs(g - gb)) < 1e-9 * max(1.0, np.max(np.abs(gb)))


def test_trifurcating_root_brute_force():
    # simudata-like multifurcation at the root with a GTR model
    t = mcp.ParseNewick("((a:0.1,b:0.2):0.05,c:0.3,(d:0.15,e:0.07):0.2);")
    ta = mcp.tree_arrays(t)
    rng = np.random.default_rng(7)
    pi, r = mcp.synthetic.random_gtr(4, rng)
    U, D, Ui, mu = mcp.GTR(pi, r)
    x = np.zeros((4, 23, ta.N), order="F")
    for v in ta.leaf_nums:
        x[rng.integers(4, size=23), np.arange(23), v - 1] = 1.0
    blv = mcp.get_branchlength_vector(t)
    rates = mcp.discrete_gamma_rates(0.5, 0.5, 4)
    ll, g = oracle.logl_grad(ta, x, U, D, Ui, mu, rates, pi, blv)
    Q, mu2 = rate_matrix("GTR", pi, r)
    llb, gb = brute_logl(ta, x, Q, mu2, rates, pi, blv, grad=True)
    assert rel_err(ll, llb) < 1e-11
    assert np.max(np.abs(g - gb)) < 1e-9 * np.max(np.abs(gb))


def test_finite_differences():
    w = mcp.synthetic.make_workload("custom", seed=11, n=12, S=4, C=50, R=4)
    blv = mcp.get_branchlength_vector(w.tree)
    ll, g = oracle.logl_grad(w.ta, w.x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, blv)
    h = 1e-6
    fd = np.zeros_like(g)
    for k in range(blv.shape[0]):
        bp, bm = blv.copy(), blv.copy()
        bp[k] += h
        bm[k] -= h
        fp = oracle.logl_grad(w.ta, w.x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, bp, grad=False)
        fm = oracle.logl_grad(w.ta, w.x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, bm, grad=False)
        fd[k] = (fp - fm) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-5 * max(1.0, np.max(np.abs(g)))


def test_pulley_principle_at_root():
    w = mcp.synthetic.make_workload("custom", seed=12, n=9, S=4, C=40, R=2)
    ta = w.ta
    blv = mcp.get_branchlength_vector(w.tree)
    kids = ta.child_idx[ta.child_ptr[ta.N - 1]:ta.child_ptr[ta.N]]
    assert kids.shape == (2,)
    a, b = kids - 1
    ll0, g0 = oracle.logl_grad(ta, w.x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, blv)
    shift = 0.3 * min(blv[a], blv[b])
    b2 = blv.copy()
    b2[a] += shift
    b2[b] -= shift
    ll1 = oracle.logl_grad(ta, w.x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, b2, grad=False)
    assert rel_err(ll1, ll0) < 1e-12
    assert abs(g0[a] - g0[b]) < 1e-9 * abs(g0[a])


def test_weights_equal_duplicated_columns():
    w = mcp.synthetic.make_workload("custom", seed=13, n=10, S=4, C=30, R=2)
    rng = np.random.default_rng(3)
    reps = rng.integers(1, 4, size=w.C)
    xd = np.asfortranarray(np.repeat(w.x, reps, axis=1))
    blv = mcp.get_branchlength_vector(w.tree)
    a = oracle.logl_grad(w.ta, w.x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, blv,
                         weights=reps.astype(float))
    b = oracle.logl_grad(w.ta, xd, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, blv)
    assert rel_err(a[0], b[0]) < 1e-12
    assert rel_err(a[1], b[1]) < 1e-10


def test_identical_categories_multiply():
    w = mcp.synthetic.make_workload("custom", seed=14, n=10, S=4, C=30, R=1)
    blv = mcp.get_branchlength_vector(w.tree)
    l1, g1 = oracle.logl_grad(w.ta, w.x, w.U, w.D, w.Uinv, w.mu, [1.0], w.pi, blv)
    l3, g3 = oracle.logl_grad(w.ta, w.x, w.U, w.D, w.Uinv, w.mu, [1.0, 1.0, 1.0], w.pi, blv)
    assert rel_err(l3, 3 * l1) < 1e-13
    assert rel_err(g3, 3 * g1) < 1e-12


def test_transition_closed_forms():
    # JC and restriction closed forms (SubstitutionModels.jl:13-50)
    U, D, Ui, mu = _jc4()
    blv = np.array([0.0, 0.05, 0.7, 3.0])
    P = oracle.transition(U, D, Ui, mu, [1.0, 0.5], blv)
    dP = oracle.transition(U, D, Ui, mu, [1.0, 0.5], blv, deriv=True)
    for r, rho in enumerate([1.0, 0.5]):
        for b, t in enumerate(blv):
            e = np.exp(-mu * rho * t)
            ref = np.full((4, 4), (1 - e) / 4)
            np.fill_diagonal(ref, 0.25 + 0.75 * e)
            assert np.max(np.abs(P[:, :, r, b] - ref)) < 1e-15
            dref = np.full((4, 4), mu * rho * e / 4)
            np.fill_diagonal(dref, -0.75 * mu * rho * e)
            assert np.max(np.abs(dP[:, :, r, b] - dref)) < 1e-14
    f = np.array([0.3, 0.7])
    U, D, Ui, mu = mcp.Restriction(f)
    P = oracle.transition(U, D, Ui, mu, [1.0], np.array([0.4]))[:, :, 0, 0]
    e = np.exp(-mu * 0.4)
    ref = f[None, :] + (np.eye(2) - f[None, :]) * e
    assert np.max(np.abs(P - ref)) < 1e-15
