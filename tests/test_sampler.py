"""NUTS harness (CPU): the ESS estimator, the host restatement of the
leapfrog with refraction, and posterior recovery on targets with known
moments.  GPU leapfrog parity lives in test_gpu_sampler.py."""

import math

import numpy as np
import pytest

from conftest import mcp

S = mcp.sampler


def test_ess_iid_and_ar1():
    rng = np.random.default_rng(0)
    x = rng.standard_normal(20000)
    assert 0.7 * 20000 < S.ess_bm(x) <= 20000
    # AR(1) with phi: integrated autocorrelation (1+phi)/(1-phi)
    phi = 0.8
    y = np.zeros(200000)
    e = rng.standard_normal(y.shape[0])
    for t in range(1, y.shape[0]):
        y[t] = phi * y[t - 1] + e[t]
    expected = y.shape[0] * (1 - phi) / (1 + phi)
    assert 0.75 * expected < S.ess_bm(y) < 1.25 * expected
    with pytest.raises(ValueError):
        S.mcse_bm(np.zeros(150))


def test_host_refraction_single_crossing_is_reflection():
    x = np.array([0.3, 1.0])
    r = np.array([-2.0, 0.5])
    g = np.zeros(2)
    xm, r1 = S._host_refraction(x, r, g, 0.25, 1e-9)
    # coordinate 0 travels 0.5: hits zero at 0.15, reflected to 0.2
    assert abs(xm[0] - 0.2) < 1e-15 and r1[0] == 2.0
    assert abs(xm[1] - 1.125) < 1e-15 and r1[1] == 0.5


def test_host_refraction_time_ordered_double_crossing():
    # two coordinates cross; each is reflected at its own crossing time and the
    # final position equals |x + eps r| coordinate-wise for straight motion
    x = np.array([0.1, 0.2, 2.0])
    r = np.array([-1.0, -1.0, 0.3])
    xm, r1 = S._host_refraction(x, r, np.zeros(3), 0.5, 1e-12)
    assert np.allclose(xm, np.abs(x + 0.5 * r), atol=1e-15)
    assert np.array_equal(r1, [1.0, 1.0, 0.3])


def test_host_refraction_molifier():
    x = np.array([0.001])
    xm, _ = S._host_refraction(x, np.zeros(1), np.zeros(1), 0.1, 0.01)
    assert abs(xm[0] - (0.001 ** 2 + 0.01 ** 2) / 0.02) < 1e-18


def _gamma_target(a, b):
    def logfgrad(x):
        return float(np.sum((a - 1) * np.log(x) - b * x)), (a - 1) / x - b
    return logfgrad


def test_nuts_recovers_gamma_moments():
    """Gamma(a=36, b=6): mean 6, variance 1, negligible mass below the
    molifier band 2*eps, where the refraction target coincides with f."""
    a, b = 36.0, 6.0
    rng = np.random.default_rng(1)
    chain = S.NUTSChain(np.full(3, 5.0), rng, target=0.7, tree_depth=8)
    draws, lfs, wall, rounds = S.run_nuts([chain], S.HostLeapfrog(_gamma_target(a, b)), 3000,
                                          n_adapt=300)
    d = draws[0]
    assert d.shape == (3000, 3)
    for k in range(3):
        se = S.mcse_bm(d[:, k])
        assert abs(d[:, k].mean() - a / b) < 4 * se
        assert abs(d[:, k].var() - a / b ** 2) < 0.15
    assert rounds > 3000


def test_nuts_with_reflections_near_zero():
    """Exponential(1) per coordinate (mode at zero): trajectories reflect at
    zero constantly; mean stays within a few MCSE of 1 up to the molifier
    band (delta = 2 eps), which only shifts mass inside [0, delta)."""
    rng = np.random.default_rng(2)
    chain = S.NUTSChain(np.full(2, 1.0), rng, target=0.8, tree_depth=8)
    lf = S.HostLeapfrog(lambda x: (float(-np.sum(x)), -np.ones_like(x)))
    draws, *_ = S.run_nuts([chain], lf, 3000, n_adapt=300)
    d = draws[0]
    assert np.all(d > 0)
    for k in range(2):
        se = S.mcse_bm(d[:, k])
        assert abs(d[:, k].mean() - 1.0) < 4 * se + chain.epsilon


def test_lockstep_chains_independent_of_batching():
    """Chains advanced together give exactly what they give alone."""
    lf = S.HostLeapfrog(_gamma_target(3.0, 1.0))
    solo = []
    for seed in (5, 6):
        c = S.NUTSChain(np.ones(2), np.random.default_rng(seed), tree_depth=5)
        d, *_ = S.run_nuts([c], lf, 30, n_adapt=10)
        solo.append(d[0])
    cs = [S.NUTSChain(np.ones(2), np.random.default_rng(seed), tree_depth=5) for seed in (5, 6)]
    both, *_ = S.run_nuts(cs, lf, 30, n_adapt=10)
    assert np.array_equal(both[0], solo[0]) and np.array_equal(both[1], solo[1])


# ------------------------------------------------------ refraction with NNI

from conftest import oracle  # noqa: E402


class _OraclePlan:
    """Plan stand-in on the CPU oracle (logl, logl_grad, set_topology)."""

    def __init__(self, w):
        self.parent = np.array(w.ta.parent, dtype=np.int32)
        self.w = w

    def set_topology(self, parent):
        self.parent = np.array(parent, dtype=np.int32)

    def logl_grad(self, blv):
        w = self.w
        ta = mcp.tree_arrays(mcp.tree_from_parent(self.parent, blv))
        return oracle.logl_grad(ta, w.x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, blv)

    def logl(self, blv):
        return self.logl_grad(blv)[0]


def _internal_nonroot(parent):
    N = len(parent)
    return [v for v in range(1, N) if np.count_nonzero(parent == v) == 2]


def test_nni_parent_is_an_interchange_and_its_own_inverse():
    w = mcp.synthetic.make_workload("custom", seed=611, n=12, S=4, C=30, R=1)
    par = np.array(w.ta.parent, dtype=np.int32)
    N = len(par)
    leaves = set(range(1, N + 1)) - set(int(p) for p in par)
    for v in _internal_nonroot(par):
        for pick in (0, 1):
            new, made = S.nni_parent(par, v, pick)
            assert made == 1 and new[N - 1] == 0
            assert np.count_nonzero(new != par) == 2          # two edges rewired
            assert set(range(1, N + 1)) - set(int(p) for p in new) == leaves
            assert all(np.count_nonzero(new == u) == 2 for u in range(1, N + 1) if u not in leaves)
            # the sibling that moved under v moves back with the same swap
            sib = int(np.flatnonzero((par == par[v - 1]) & (np.arange(1, N + 1) != v))[0]) + 1
            back_pick = int(np.flatnonzero(np.flatnonzero(new == v) + 1 == sib)[0])
            again, made2 = S.nni_parent(new, v, back_pick)
            assert made2 == 1 and np.array_equal(again, par)
    assert S.nni_parent(par, min(leaves), 0)[1] == 0        # a leaf
    assert S.nni_parent(par, N, 0)[1] == 0                  # the root


def test_nni_refraction_without_crossing_is_the_plain_leapfrog():
    w = mcp.synthetic.make_workload("custom", seed=612, n=10, S=4, C=40, R=1)
    blv = mcp.get_branchlength_vector(w.tree)
    plan = _OraclePlan(w)
    rng = np.random.default_rng(3)
    r = 0.05 * rng.standard_normal(blv.shape)
    x = S.TreePosition(blv, plan.parent)
    lf0, g0 = plan.logl_grad(blv)
    lf_ = S.NNIRefractionLeapfrog(plan)
    s = lf_.one(x, r, g0, 1e-4, 1e-300)
    ref = S.HostLeapfrog(plan.logl_grad).one(blv, r, g0, 1e-4, 1e-300)
    # positions do not depend on the new gradient; logL and momenta agree to
    # the oracle's threaded summation (two calls may differ in the last bit)
    assert np.array_equal(np.asarray(s.x), ref.x)
    assert np.max(np.abs(s.r - ref.r)) <= 1e-13 * np.max(np.abs(ref.r))
    assert abs(s.lf - ref.lf) <= 1e-13 * abs(ref.lf) and np.array_equal(s.x.parent, plan.parent)
    assert lf_.att_nni == 0


@pytest.mark.parametrize("accept", [True, False])
def test_nni_attempt_follows_the_momentum_rule(accept):
    """One internal branch reaches zero (the only moving coordinate): the
    interchange is kept iff r^2 > 2 (U_before - U_after), and then the
    branch's momentum becomes +sqrt(r^2 - 2 dU) (refraction.jl:56-87)."""
    w = mcp.synthetic.make_workload("custom", seed=613, n=10, S=4, C=200, R=1)
    blv = mcp.get_branchlength_vector(w.tree)
    par0 = np.array(w.ta.parent, dtype=np.int32)
    plan = _OraclePlan(w)
    delta = 1e-3
    pick = int(np.random.default_rng(7).integers(2))
    k = None
    for v in _internal_nonroot(par0):
        cand, _ = S.nni_parent(par0, v, pick)
        xm = blv.copy()
        xm[v - 1] = delta / 2.0                # the molified reflection point
        plan.set_topology(par0)
        ub = plan.logl(xm)
        plan.set_topology(cand)
        ua = plan.logl(xm)
        if ub - ua > 0.5:                      # the interchange costs energy
            k, d_u = v - 1, 2.0 * (ub - ua)
            break
    assert k is not None
    plan.set_topology(par0)
    a = blv[k]
    vel = math.sqrt(d_u) * (1.5 if accept else 0.5)
    r = np.zeros_like(blv)
    r[k] = -vel
    eps = 1.5 * a / vel
    lf_ = S.NNIRefractionLeapfrog(plan, rng=np.random.default_rng(7))
    s = lf_.one(S.TreePosition(blv, par0), r, np.zeros_like(blv), eps, delta)
    assert lf_.att_nni == 1
    if accept:
        assert lf_.nni == 1 and not np.array_equal(s.x.parent, par0)
        r_k = math.sqrt(vel * vel - d_u)
    else:
        assert lf_.nni == 0 and np.array_equal(s.x.parent, par0)
        r_k = vel                              # plain reflection
    # the branch restarts from zero and travels the rest of the step
    y = r_k * (eps - a / vel)
    exp_x = y if y >= delta else (y * y + delta * delta) / (2 * delta)
    assert abs(s.x[k] - exp_x) <= 1e-12 * max(1.0, exp_x)
    plan.set_topology(s.x.parent)
    lf, g = plan.logl_grad(np.asarray(s.x))
    assert abs(s.lf - lf) <= 1e-13 * abs(lf)
