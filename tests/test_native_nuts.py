"""CPU: the native lockstep NUTS driver (mcp_nuts_*) with a host leapfrog
callback -- posterior recovery on targets with known moments, determinism,
and that lockstep batching leaves each chain's trajectory unchanged."""

import numpy as np

from conftest import mcp

S = mcp.sampler


def _callback(logfgrad, dims, prior=None):
    host = S.HostLeapfrog(logfgrad, prior=prior)
    off = np.concatenate([[0], np.cumsum(dims)])

    def lf(ed, x, r, g, lp):
        for i in range(len(dims)):
            a, b = off[i], off[i + 1]
            st = host.one(x[a:b].copy(), r[a:b].copy(), g[a:b].copy(), ed[2 * i], ed[2 * i + 1])
            x[a:b] = st.x
            r[a:b] = st.r
            g[a:b] = st.g
            lp[i] = st.lf
    return lf


def _gamma(a, b):
    def f(x):
        return float(np.sum((a - 1) * np.log(x) - b * x)), (a - 1) / x - b
    return f


def test_native_nuts_recovers_gamma_moments():
    a, b = 36.0, 6.0
    dims = [3, 3]
    nuts = S.NativeNUTS([np.full(3, 5.0), np.full(3, 7.0)], [1, 2], target=0.7, tree_depth=8)
    nuts.use_callback(_callback(_gamma(a, b), dims))
    nuts.run(300, adapt=True, keep_draws=False)
    draws, lp, wall, rounds, lfs = nuts.run(2000)
    assert lp.shape == (2000, 2) and np.all(np.isfinite(lp))
    for d in draws:
        for k in range(3):
            se = S.mcse_bm(d[:, k])
            assert abs(d[:, k].mean() - a / b) < 4.5 * se
            assert abs(d[:, k].var() - a / b ** 2) < 0.15
    assert lfs >= rounds


def test_native_nuts_reflection_exponential_exact():
    """Exponential(1) in 4 coordinates (mass at zero): with the molifier off
    the sampler is exact -- mean 1, variance 1."""
    nuts = S.NativeNUTS([np.full(4, 1.0)], [5], target=0.8, tree_depth=8, molifier=False)

    def f(x):
        return float(-np.sum(x)), -np.ones_like(x)
    nuts.use_callback(_callback(f, [4]))
    nuts.run(300, adapt=True, keep_draws=False)
    draws, *_ = nuts.run(4000)
    d = draws[0]
    for k in range(4):
        se = S.mcse_bm(d[:, k])
        assert abs(d[:, k].mean() - 1.0) < 4.5 * se
        sq = (d[:, k] - 1.0) ** 2          # E = variance = 1
        assert abs(sq.mean() - 1.0) < 4.5 * S.mcse_bm(sq)
    assert np.all(d > 0)


def test_native_nuts_deterministic_and_lockstep_invariant():
    f = _gamma(9.0, 3.0)
    x0 = [np.full(3, 2.0), np.full(5, 4.0), np.full(2, 1.0)]

    def run(xs, seeds, n):
        nuts = S.NativeNUTS(xs, seeds, target=0.8, tree_depth=6)
        nuts.use_callback(_callback(f, [len(v) for v in xs]))
        nuts.run(20, adapt=True, keep_draws=False)
        draws, lp, *_ = nuts.run(n)
        return draws, lp, nuts.state()
    d1, l1, s1 = run(x0, [11, 12, 13], 50)
    d2, l2, s2 = run(x0, [11, 12, 13], 50)
    assert all(np.array_equal(a, b) for a, b in zip(d1, d2)) and np.array_equal(l1, l2)
    # chain 1 alone follows exactly the trajectory it follows in the batch
    d3, l3, _ = run([x0[1]], [12], 50)
    assert np.array_equal(d3[0], d1[1]) and np.array_equal(l3[:, 0], l1[:, 1])
    assert s1[0].shape == (10,) and np.all(s1[1] > 0)


def test_native_and_host_samplers_agree_in_distribution():
    a, b = 20.0, 4.0
    rng = np.random.default_rng(3)
    chain = S.NUTSChain(np.full(2, 5.0), rng, target=0.8, tree_depth=8)
    dh, *_ = S.run_nuts([chain], S.HostLeapfrog(_gamma(a, b)), 3000, n_adapt=300)
    nuts = S.NativeNUTS([np.full(2, 5.0)], [4], target=0.8, tree_depth=8)
    nuts.use_callback(_callback(_gamma(a, b), [2]))
    nuts.run(300, adapt=True, keep_draws=False)
    dn, *_ = nuts.run(3000)
    for k in range(2):
        se = np.hypot(S.mcse_bm(dh[0][:, k]), S.mcse_bm(dn[0][:, k]))
        assert abs(dh[0][:, k].mean() - dn[0][:, k].mean()) < 4.5 * se
