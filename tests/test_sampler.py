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
