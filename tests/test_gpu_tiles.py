"""GPU: every built tile width (columns per thread) of the level and fused
kernels against the CPU oracle, and the autotuner on a mixed batch."""

import numpy as np
import pytest

from conftest import mcp, oracle, rel_err
from test_gpu_parity import GRAD_TOL, LL_TOL, grad_err

pytestmark = pytest.mark.gpu

CANDIDATES = {2: [1, 2, 4], 3: [1, 2, 4], 4: [1, 2, 4], 20: [0, 1]}


def _case(seed, n, S, C, R):
    w = mcp.synthetic.make_workload("custom", seed=seed, n=n, S=S, C=C, R=R)
    blv = mcp.get_branchlength_vector(w.tree)
    llo, go = oracle.logl_grad(w.ta, w.x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, blv)
    return w, blv, llo, go


@pytest.mark.parametrize("seed,n,S,C,R", [(501, 23, 2, 700, 4), (502, 19, 3, 300, 2),
                                          (503, 31, 4, 1100, 4), (504, 12, 20, 300, 2)])
@pytest.mark.parametrize("schedule", ["levels", "fused"])
def test_every_tile_width_matches_oracle(seed, n, S, C, R, schedule):
    print(f"tiles seed={seed} S={S} C={C} R={R} {schedule}", flush=True)
    w, blv, llo, go = _case(seed, n, S, C, R)
    p = mcp.Plan(w.ta.parent, w.x, R)
    p.set_model(mcp.Model(w.U, w.D, w.Uinv, w.mu, w.rates, w.pi))
    p.set_schedule(schedule)
    for cpt in CANDIDATES[S]:
        p.set_tiles(S, cpt)
        assert p.tiles(S) == cpt
        ll, g = p.logl_grad(blv)
        assert rel_err(ll, llo) < LL_TOL, cpt
        assert grad_err(g, go) < GRAD_TOL, cpt
        assert rel_err(p.logl(blv), llo) < LL_TOL, cpt
    p.set_tiles(S, -1)
    with pytest.raises(mcp.McpArgumentError):
        p.set_tiles(S, 3)
    p.close()


def test_autotune_mixed_batch_keeps_results_and_reports_candidates():
    data = mcp.synthetic.make_batched_workload(n_trees=6, chains=2, n_range=(16, 60),
                                               c_range=(300, 2000))
    ds, items, models, blv = data
    b = mcp.Batch(ds, items)
    for i, m in enumerate(models):
        b.set_model(i, mcp.Model(*m))
    ll0, g0 = b.logl_grad(blv)
    chosen, rows = b.autotune(reps=2)
    S_present = sorted({b.item_S[i] for i in range(len(items))})
    assert sorted(chosen) == S_present
    assert len(rows) == sum(len(CANDIDATES[S]) for S in S_present)
    for S, cpt in chosen.items():
        best = min((r for r in rows if r[0] == S), key=lambda r: r[2])
        assert cpt == best[1]
    ll1, g1 = b.logl_grad(blv)
    assert np.max(np.abs(ll1 - ll0) / np.abs(ll0)) < 1e-12
    assert np.max(np.abs(g1 - g0)) <= 1e-10 * np.max(np.abs(g0))
    # a stored choice reproduces the tuned results bit for bit in a new engine
    b2 = mcp.Batch(ds, items)
    for i, m in enumerate(models):
        b2.set_model(i, mcp.Model(*m))
    for S, cpt in chosen.items():
        b2.set_tiles(S, cpt)
    ll2, g2 = b2.logl_grad(blv)
    assert np.array_equal(ll1, ll2) and np.array_equal(g1, g2)
    # graph replay after tuning
    b.use_graph(True)
    ll3, g3 = b.logl_grad(blv)
    assert np.array_equal(ll3, ll1) and np.array_equal(g3, g1)
    b.close()
    b2.close()


@pytest.mark.parametrize("width", [1, 2, 4, 8])
def test_hybrid_and_fused_schedules_match_levels(width):
    data = mcp.synthetic.make_batched_workload(n_trees=5, chains=2, n_range=(16, 80),
                                               c_range=(300, 3000))
    ds, items, models, blv = data
    res = {}
    for mode in ("levels", "hybrid", "fused"):
        b = mcp.Batch(ds, items)
        for i, m in enumerate(models):
            b.set_model(i, mcp.Model(*m))
        b.set_schedule(mode, hybrid_width=width)
        assert b.schedule()[0] == mode
        ll, g = b.logl_grad(blv)
        l_only = b.logl(blv)
        b.set_timing(True)
        b.logl_grad(blv)
        ms, ph = b.launch_times()
        by = b.launch_bytes()
        b.set_timing(False)
        fwd, adj = b.algorithmic_bytes()
        assert by.shape == ms.shape
        assert abs(by.sum() - (fwd + adj)) <= 1e-9 * (fwd + adj)
        res[mode] = (ll, g, l_only)
        b.close()
    for mode in ("hybrid", "fused"):
        # A fused launch folds each tile's log scalers into one product per
        # lane (one log per pass instead of one per node), so logL agrees
        # with the levels schedule to rounding.  The gradient sums are the
        # same operations in the same order: bit for bit.
        np.testing.assert_allclose(res[mode][0], res["levels"][0], rtol=1e-13, atol=0)
        assert np.array_equal(res[mode][1], res["levels"][1]), mode
        np.testing.assert_allclose(res[mode][2], res["levels"][2], rtol=1e-13, atol=0)


def test_hybrid_schedule_single_tree_vs_oracle_and_schedule_autotune():
    w, blv, llo, go = _case(505, 64, 4, 900, 4)
    p = mcp.Plan(w.ta.parent, w.x, 4)
    p.set_model(mcp.Model(w.U, w.D, w.Uinv, w.mu, w.rates, w.pi))
    p.set_schedule("hybrid", hybrid_width=3)
    ll, g = p.logl_grad(blv)
    assert rel_err(ll, llo) < LL_TOL and grad_err(g, go) < GRAD_TOL
    (mode, width), rows = p.autotune_schedule(reps=2)
    assert len(rows) == 7
    best = min(rows, key=lambda r: r[2])
    assert mode == best[0] and (mode != "hybrid" or width == best[1])
    ll2, g2 = p.logl_grad(blv)
    # the tuned schedule may differ from hybrid width 3: logL to rounding
    # (fused launches sum log scalers as one product per lane), gradient exact
    assert rel_err(ll2, ll) < 1e-13 and np.array_equal(g2, g)
    p.close()


@pytest.mark.parametrize("cpt", [1, 4])
def test_fused_tree_too_large_for_the_lds_code_stage(cpt):
    """600 leaves: 600 * 64 * CPT bytes of tip codes exceed the 32 KB LDS
    stage, so the fused launch's leaves read their codes from global memory.
    Same results as the levels schedule (gradient bit for bit) and the oracle."""
    print(f"fused without LDS stage cpt={cpt}", flush=True)
    w, blv, llo, go = _case(511, 600, 4, 150, 2)
    res = {}
    for sched in ("levels", "fused"):
        p = mcp.Plan(w.ta.parent, w.x, 2)
        p.set_model(mcp.Model(w.U, w.D, w.Uinv, w.mu, w.rates, w.pi))
        p.set_tiles(4, cpt)
        p.set_schedule(sched)
        res[sched] = p.logl_grad(blv)
        p.close()
    ll, g = res["fused"]
    assert rel_err(ll, llo) < LL_TOL and grad_err(g, go) < GRAD_TOL
    assert rel_err(ll, res["levels"][0]) < 1e-13 and np.array_equal(g, res["levels"][1])


def test_fused_schedule_for_every_wave_count():
    """The fused schedule groups each item's nodes into Hu steps of at most
    one node per wavefront of a block, and lays the pre-order slabs out for
    that order; every block size gives the oracle's results (gradient bit for
    bit the levels schedule's)."""
    w, blv, llo, go = _case(521, 45, 4, 400, 3)
    p = mcp.Plan(w.ta.parent, w.x, 3)
    p.set_model(mcp.Model(w.U, w.D, w.Uinv, w.mu, w.rates, w.pi))
    p.set_schedule("levels")
    ll_l, g_l = p.logl_grad(blv)
    p.set_schedule("fused")
    for waves in (1, 2, 3, 8, 4):
        print(f"fused waves={waves}", flush=True)
        p.set_fused_waves(waves)
        ll, g = p.logl_grad(blv)
        assert rel_err(ll, llo) < LL_TOL and grad_err(g, go) < GRAD_TOL, waves
        assert rel_err(ll, ll_l) < 1e-13 and np.array_equal(g, g_l), waves
        ll_m, g_m, dP, dpi = p.logl_grad_model(blv)
        assert np.array_equal(g_m, g_l), waves
    p.close()


def _fused_runs(make, monkeypatch):
    """(logL, gradient, model gradient, algorithmic bytes) of a fused-schedule
    engine built by make(), with the LDS slab stage on, then off."""
    out = []
    for off in (False, True):
        if off:
            monkeypatch.setenv("MCP_FUSED_LDS_SLABS", "0")
        else:
            monkeypatch.delenv("MCP_FUSED_LDS_SLABS", raising=False)
        eng, blv = make()
        eng.set_schedule("fused")
        ll, g = eng.logl_grad(blv)
        mg = eng.logl_grad_model(blv) if hasattr(eng, "logl_grad_model") else None
        out.append((ll, g, mg, eng.algorithmic_bytes()))
        eng.close()
    monkeypatch.delenv("MCP_FUSED_LDS_SLABS", raising=False)
    return out


@pytest.mark.parametrize("case", ["coalescent40", "caterpillar30", "trifurcating", "cpt2"])
def test_fused_lds_slab_stage_matches_hbm_slabs(case, monkeypatch):
    """Under the fused schedule a block keeps the post-order frontier's
    partials and the pre-order slabs in LDS when they fit.  The arithmetic is
    the same as with every slab in HBM: logL, gradient and model gradient bit
    for bit, and the algorithmic bytes drop by one partial read (post-order)
    and one pre-order slab write and read per internal non-root node."""
    print(f"lds slab stage {case}", flush=True)
    if case == "trifurcating":
        t = mcp.ParseNewick("(((A:0.2,B:0.1):0.05,(C:0.3,D:0.01):0.2):0.1,"
                            "((E:0.6,F:0.02):0.3,G:0.1):0.2,H:0.4);")
        ta = mcp.tree_arrays(t)
        rng = np.random.default_rng(7)
        C, R = 300, 3
        x = np.zeros((4, C, ta.N), order="F")
        for leaf in ta.leaf_nums:
            x[rng.integers(0, 4, C), np.arange(C), leaf - 1] = 1.0
        pi = rng.dirichlet(np.full(4, 5.0))
        U, D, Uinv, mu = mcp.GTR(pi, rng.exponential(1.0, 6))
        model = mcp.Model(U, D, Uinv, mu, np.asarray(mcp.discrete_gamma_rates(0.5, 0.5, R)), pi)
        parent, blv, cpt = ta.parent, mcp.get_branchlength_vector(t), 1
        llo, go = oracle.logl_grad(ta, x, U, D, Uinv, mu, model.rates, pi, blv)
    else:
        n, S, topo, cpt = {"coalescent40": (40, 4, "coalescent", 1),
                           "caterpillar30": (30, 2, "caterpillar", 1),
                           "cpt2": (25, 4, "yule", 2)}[case]
        R = 4
        w = mcp.synthetic.make_workload("custom", seed=531, n=n, S=S, C=700, R=R, topology=topo)
        x, parent, blv = w.x, w.ta.parent, mcp.get_branchlength_vector(w.tree)
        model = mcp.Model(w.U, w.D, w.Uinv, w.mu, w.rates, w.pi)
        llo, go = oracle.logl_grad(w.ta, w.x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, blv)
    S = x.shape[0]

    def make():
        p = mcp.Plan(parent, x, R)
        p.set_model(model)
        p.set_tiles(S, cpt)
        return p, blv

    (ll_on, g_on, mg_on, b_on), (ll_off, g_off, mg_off, b_off) = _fused_runs(make, monkeypatch)
    assert rel_err(ll_on, llo) < LL_TOL and grad_err(g_on, go) < GRAD_TOL
    assert ll_on == ll_off and np.array_equal(g_on, g_off)
    for a, b in zip(mg_on, mg_off):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    d_fwd, d_adj = b_off[0] - b_on[0], b_off[1] - b_on[1]
    assert d_fwd > 0 and d_adj == 2 * d_fwd


def test_fused_lds_slab_stage_in_a_mixed_batch(monkeypatch):
    """A batch where the slab stage fits the small trees' blocks and not the
    large ones' (CPT 4: 8 KB slots): both kinds of item in one launch give
    the results of the all-HBM form bit for bit."""
    data = mcp.synthetic.make_batched_workload(n_trees=5, chains=2, n_range=(8, 200),
                                               c_range=(300, 1500), seed=20261209)
    ds, items, models, blv = data

    def make():
        b = mcp.Batch(ds, items)
        for i, m in enumerate(models):
            b.set_model(i, mcp.Model(*m))
        for S in sorted({b.item_S[i] for i in range(len(items))}):
            b.set_tiles(S, 4)
        return b, blv

    (ll_on, g_on, _, b_on), (ll_off, g_off, _, b_off) = _fused_runs(make, monkeypatch)
    assert np.array_equal(ll_on, ll_off) and np.array_equal(g_on, g_off)
    assert b_on[0] < b_off[0] and b_on[1] < b_off[1]
