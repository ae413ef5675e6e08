"""GPU parity: the HIP engine (through the C ABI) against the reference's
golden values and the CPU oracle.

Tolerances (from the accuracy target of this path):
  logL      |ll_gpu - ll_ref| <= 1e-10 * |ll_ref|
  gradient  |g_gpu - g_ref|_k <= 1e-8 * max(|g_ref_k|, 1e-3 * max_j |g_ref_j|)
            (relative per entry, with a floor for entries that are small
            because of cancellation between sites)
"""

import os

import numpy as np
import pytest

from conftest import PRIMATES_LL, SIMU_GRAD, load_fixture, mcp, oracle, rel_err

pytestmark = pytest.mark.gpu

LL_TOL = 1e-10
GRAD_TOL = 1e-8


def grad_err(g, ref):
    ref = np.asarray(ref)
    floor = 1e-3 * np.max(np.abs(ref)) if ref.size else 0.0
    return float(np.max(np.abs(np.asarray(g) - ref) / np.maximum(np.abs(ref), floor)))


def _oracle(w, x=None, weights=None, blv=None, grad=True):
    blv = mcp.get_branchlength_vector(w.tree) if blv is None else blv
    return oracle.logl_grad(w.ta, w.x if x is None else x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi,
                            blv, weights=weights if weights is not None else w.weights, grad=grad)


def _plan(w, schedule="fused", **kw):
    p = mcp.Plan(w.ta.parent, w.x, len(w.rates), weights=w.weights, **kw)
    p.set_model(mcp.Model(w.U, w.D, w.Uinv, w.mu, w.rates, w.pi))
    p.set_schedule(schedule)
    return p


SCHEDULES = ["fused", "levels"]


def test_primates_golden_phylodist():
    t, x = load_fixture("primates")
    pd = mcp.PhyloDist(t, np.ones(4) / 4, [1.0], [1.0], mcp.JC)
    assert mcp.size(pd) == (4, 1, 22)
    ll = mcp.logpdf(pd, x)
    assert rel_err(ll, PRIMATES_LL) < LL_TOL


def test_simudata_gradient_golden_phylodist():
    t, x = load_fixture("simudata")
    pd = mcp.PhyloDist(t, np.ones(4) / 4, [1.0], [1.0], mcp.JC)
    ll, g = mcp.gradlogpdf(pd, x)
    U, D, Ui, mu = mcp.JC(np.ones(4) / 4)
    llo, go = oracle.logl_grad(mcp.tree_arrays(t), x, U, D, Ui, mu, [1.0], np.ones(4) / 4,
                               mcp.get_branchlength_vector(t))
    assert rel_err(ll, llo) < LL_TOL
    assert grad_err(g, SIMU_GRAD) < GRAD_TOL
    assert grad_err(g, go) < GRAD_TOL


@pytest.mark.parametrize("compress", [False, True])
def test_uncompressed_vs_compressed_primates(compress):
    t, x = load_fixture("primates")
    pd = mcp.PhyloDist(t, np.ones(4) / 4, [1.0], [1.0], mcp.JC, compress=compress)
    assert rel_err(mcp.logpdf(pd, x), PRIMATES_LL) < LL_TOL


CASES = [
    # (seed, n, S, C, R, topology, gaps)
    (101, 8, 2, 300, 1, "yule", False),
    (102, 33, 4, 1000, 4, "coalescent", True),
    (103, 17, 3, 257, 2, "caterpillar", True),
    (104, 40, 20, 700, 4, "coalescent", True),
    (105, 64, 4, 129, 1, "caterpillar", False),
    (106, 25, 5, 500, 3, "yule", True),
    (107, 12, 16, 333, 2, "coalescent", False),
]


@pytest.mark.parametrize("schedule", SCHEDULES)
@pytest.mark.parametrize("seed,n,S,C,R,topo,gaps", CASES)
def test_random_cases_vs_oracle(seed, n, S, C, R, topo, gaps, schedule):
    print(f"case seed={seed} n={n} S={S} C={C} R={R} {topo} {schedule}", flush=True)
    w = mcp.synthetic.make_workload("custom", seed=seed, n=n, S=S, C=C, R=R, topology=topo)
    x = np.array(w.x, order="F")
    if gaps:
        rng = np.random.default_rng(seed)
        for v in w.ta.leaf_nums:
            cols = rng.choice(C, size=C // 10, replace=False)
            x[:, cols, v - 1] = 1.0
    w.x = x
    llo, go = _oracle(w)
    p = _plan(w, schedule)
    ll, g = p.logl_grad(mcp.get_branchlength_vector(w.tree))
    assert rel_err(ll, llo) < LL_TOL
    assert grad_err(g, go) < GRAD_TOL
    assert rel_err(p.logl(mcp.get_branchlength_vector(w.tree)), llo) < LL_TOL


@pytest.mark.parametrize("schedule", SCHEDULES)
def test_gtr100_full_size_vs_oracle_and_fd(schedule):
    """The 100-taxon 4-state GTR workload: 10,000 weighted patterns, 4 gamma
    categories; oracle parity and central finite differences (h = 1e-6)."""
    w = mcp.synthetic.make_workload("gtr100")
    assert w.C == 10_000 and w.N == 199
    blv = mcp.get_branchlength_vector(w.tree)
    llo, go = _oracle(w)
    p = _plan(w, schedule)
    ll, g = p.logl_grad(blv)
    assert rel_err(ll, llo) < LL_TOL
    assert grad_err(g, go) < GRAD_TOL
    h = 1e-6
    for k in [0, 5, 50, 99, 100, 150, 197]:
        bp, bm = blv.copy(), blv.copy()
        bp[k] += h
        bm[k] -= h
        fd = (p.logl(bp) - p.logl(bm)) / (2 * h)
        assert abs(fd - g[k]) < 2e-5 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_multiple_phylodist_primates():
    t, x = load_fixture("primates")
    freqs = np.full((4, 2), 0.25)
    mpd = mcp.MultiplePhyloDist([t, t], freqs, [1.0], [1.0], mcp.JC)
    assert mcp.size(mpd) == (4, 1, 22, 2)
    mdf = np.stack([x, x], axis=3)
    assert rel_err(mcp.logpdf(mpd, mdf), 2 * PRIMATES_LL) < LL_TOL
    res = mcp.__logpdf(mpd, mdf)
    assert rel_err(res[0][0], PRIMATES_LL) < LL_TOL
    assert rel_err(res[1][0], PRIMATES_LL) < LL_TOL
    with pytest.raises(ValueError):
        mcp.MultiplePhyloDist([t, t], np.full((4, 3), 0.25), [1.0], [1.0], mcp.JC)


@pytest.mark.parametrize("schedule", SCHEDULES)
def test_ragged_batch_vs_oracle(schedule):
    specs = [(201, 10, 2, 300, 4), (202, 40, 4, 1500, 4), (203, 25, 4, 640, 1),
             (204, 16, 2, 70, 2), (205, 31, 20, 200, 4)]
    ws = [mcp.synthetic.make_workload("custom", seed=s, n=n, S=S, C=C, R=R)
          for s, n, S, C, R in specs]
    datasets, items = [], []
    for i, w in enumerate(ws):
        leaf = w.ta.leaf_nums
        datasets.append({"x_leaves": w.x[:, :, leaf - 1], "leaf_nums": leaf})
        items.append({"dataset": i, "parent": w.ta.parent, "n_rates": len(w.rates)})
    # a second chain on dataset 1 with perturbed branch lengths
    items.append({"dataset": 1, "parent": ws[1].ta.parent, "n_rates": 4})
    b = mcp.Batch(datasets, items)
    b.set_schedule(schedule)
    for i, w in enumerate(ws):
        b.set_model(i, mcp.Model(w.U, w.D, w.Uinv, w.mu, w.rates, w.pi))
    w1 = ws[1]
    b.set_model(len(ws), mcp.Model(w1.U, w1.D, w1.Uinv, w1.mu, w1.rates, w1.pi))
    blvs = [mcp.get_branchlength_vector(w.tree) for w in ws]
    extra = blvs[1] * np.exp(np.random.default_rng(9).normal(0, 0.1, blvs[1].shape))
    ll, g = b.logl_grad(np.concatenate(blvs + [extra]))
    gs = b.split(g)
    for i, w in enumerate(ws):
        llo, go = _oracle(w)
        assert rel_err(ll[i], llo) < LL_TOL, i
        assert grad_err(gs[i], go) < GRAD_TOL, i
    llo, go = _oracle(w1, blv=extra)
    assert rel_err(ll[-1], llo) < LL_TOL
    assert grad_err(gs[-1], go) < GRAD_TOL
    assert np.allclose(b.logl(np.concatenate(blvs + [extra])), ll, rtol=1e-14, atol=0)


def _nni(ta):
    """Swap a grandchild across an internal edge: returns a new parent array."""
    N = ta.N
    par = ta.parent.copy()
    nch = np.diff(ta.child_ptr)
    for v in range(1, N):          # internal non-root v
        if nch[v - 1] == 0 or par[v - 1] == 0:
            continue
        p = par[v - 1]
        sib = [u for u in ta.child_idx[ta.child_ptr[p - 1]:ta.child_ptr[p]] if u != v][0]
        kid = ta.child_idx[ta.child_ptr[v - 1]]
        par[kid - 1], par[sib - 1] = p, v
        return par
    raise AssertionError("no internal edge")


@pytest.mark.parametrize("schedule", SCHEDULES)
def test_topology_change_vs_oracle(schedule):
    w = mcp.synthetic.make_workload("custom", seed=301, n=30, S=4, C=400, R=2)
    p = _plan(w, schedule)
    blv = mcp.get_branchlength_vector(w.tree)
    par2 = _nni(w.ta)
    p.set_topology(par2)
    t2 = mcp.tree_from_parent(par2, blv)
    ta2 = mcp.tree_arrays(t2)
    llo, go = oracle.logl_grad(ta2, w.x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, blv)
    ll, g = p.logl_grad(blv)
    assert rel_err(ll, llo) < LL_TOL
    assert grad_err(g, go) < GRAD_TOL


@pytest.mark.parametrize("schedule", SCHEDULES)
def test_graph_replays_match_eager(schedule):
    w = mcp.synthetic.make_workload("custom", seed=401, n=50, S=4, C=2000, R=4)
    p = _plan(w, schedule)
    rng = np.random.default_rng(5)
    blv0 = mcp.get_branchlength_vector(w.tree)
    blvs = [blv0 * np.exp(rng.normal(0, 0.2, blv0.shape)) for _ in range(4)]
    eager = [p.logl_grad(b) for b in blvs]
    p.use_graph(True)
    for rep in range(2):
        for b, (le, ge) in zip(blvs, eager):
            lg, gg = p.logl_grad(b)
            assert lg == le and np.array_equal(gg, ge)
            assert p.logl(b) == le
    p.use_graph(False)


def test_deterministic_and_schedules_agree():
    w = mcp.synthetic.make_workload("custom", seed=402, n=60, S=4, C=3000, R=4)
    p = _plan(w)
    b = mcp.get_branchlength_vector(w.tree)
    a1 = p.logl_grad(b)
    a2 = p.logl_grad(b)
    assert a1[0] == a2[0] and np.array_equal(a1[1], a2[1])
    # the two schedules run the same tasks; the gradient reductions are the
    # same, the fused log-scaler sum is one product per lane (rounding only)
    p.set_schedule("levels")
    a3 = p.logl_grad(b)
    assert rel_err(a3[0], a1[0]) < 1e-13 and np.array_equal(a3[1], a1[1])


def test_errors():
    w = mcp.synthetic.make_workload("custom", seed=403, n=6, S=4, C=20, R=1)
    with pytest.raises(mcp.McpDimensionMismatch):
        mcp.Plan(w.ta.parent[:-1], w.x, 1)
    p = mcp.Plan(w.ta.parent, w.x, 1)
    with pytest.raises(mcp.McpError):
        p.logl(mcp.get_branchlength_vector(w.tree))   # model not set
    with pytest.raises(mcp.McpDimensionMismatch):
        p.set_model(mcp.Model(w.U, w.D, w.Uinv, w.mu, [1.0, 2.0], w.pi))
    bad = w.ta.parent.copy()
    bad[0] = bad[0] + 100
    with pytest.raises(mcp.McpArgumentError):
        p.set_topology(bad)
    x9 = np.zeros((9, 20, w.N), order="F")
    x9[0] = 1.0
    with pytest.raises(mcp.McpArgumentError):
        mcp.Plan(w.ta.parent, x9, 1)                    # no kernel for 9 states


def test_numeric_failure_propagates_as_values():
    """log 0 and NaN branch lengths come back as values, as in the reference."""
    w = mcp.synthetic.make_workload("custom", seed=404, n=6, S=4, C=20, R=1)
    x = np.array(w.x, order="F")
    x[:, 3, w.ta.leaf_nums[0] - 1] = 0.0              # impossible column
    w.x = x
    p = _plan(w)
    blv = mcp.get_branchlength_vector(w.tree)
    assert np.isneginf(p.logl(blv)) or np.isnan(p.logl(blv))
    blv[2] = np.nan
    assert np.isnan(p.logl(blv))


def test_rate_gradient_components_and_fd():
    """Per-(branch, category) gradient components sum to the branch gradient;
    d logL / d rho_r (and, through the discrete gamma rates, d logL / d alpha)
    match central finite differences of the CPU oracle."""
    w = mcp.synthetic.make_workload("custom", seed=601, n=25, S=4, C=600, R=4)
    blv = mcp.get_branchlength_vector(w.tree)
    p = _plan(w, "levels")
    ll, G = p.logl_grad_by_rate(blv)
    ll2, g = p.logl_grad(blv)
    assert G.shape == (w.N - 1, 4)
    assert rel_err(ll, ll2) < 1e-14
    assert np.max(np.abs(G.sum(axis=1) - g)) <= 1e-12 * np.max(np.abs(g))
    _, dr = p.logl_grad_rates(blv, w.rates)
    base = np.array(w.rates, dtype=np.float64)
    for r in range(4):
        h = 1e-6 * base[r]
        rp, rm = base.copy(), base.copy()
        rp[r] += h
        rm[r] -= h
        fp = oracle.logl_grad(w.ta, w.x, w.U, w.D, w.Uinv, w.mu, rp, w.pi, blv, grad=False)
        fm = oracle.logl_grad(w.ta, w.x, w.U, w.D, w.Uinv, w.mu, rm, w.pi, blv, grad=False)
        fd = (fp - fm) / (2 * h)
        assert abs(fd - dr[r]) <= 1e-6 * max(1.0, abs(dr[r])), (r, fd, dr[r])
    # gamma shape alpha (rates = discrete_gamma_rates(alpha, alpha, 4))
    a, ha = 0.5, 1e-6
    drho = (np.asarray(mcp.discrete_gamma_rates(a + ha, a + ha, 4))
            - np.asarray(mcp.discrete_gamma_rates(a - ha, a - ha, 4))) / (2 * ha)
    da = float(dr @ drho)
    fp = oracle.logl_grad(w.ta, w.x, w.U, w.D, w.Uinv, w.mu,
                          mcp.discrete_gamma_rates(a + ha, a + ha, 4), w.pi, blv, grad=False)
    fm = oracle.logl_grad(w.ta, w.x, w.U, w.D, w.Uinv, w.mu,
                          mcp.discrete_gamma_rates(a - ha, a - ha, 4), w.pi, blv, grad=False)
    assert abs((fp - fm) / (2 * ha) - da) <= 1e-5 * max(1.0, abs(da))
    # batch form: same components per item
    b = mcp.Batch([{"x_leaves": w.x[:, :, w.ta.leaf_nums - 1], "leaf_nums": w.ta.leaf_nums}],
                  [{"dataset": 0, "parent": w.ta.parent, "n_rates": 4}] * 2)
    for i in range(2):
        b.set_model(i, mcp.Model(w.U, w.D, w.Uinv, w.mu, w.rates, w.pi))
    llb, Gb = b.logl_grad_by_rate(np.concatenate([blv, blv]))
    assert np.array_equal(Gb[0], Gb[1]) and np.max(np.abs(Gb[0] - G)) <= 1e-12 * np.max(np.abs(G))
