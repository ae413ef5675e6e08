"""GPU: edge shapes of the pruning path through the C ABI against the CPU
oracle -- a root with one leaf, a cherry, a star (all leaves under a
multifurcating root), a zero-length internal branch, extreme branch lengths
(1e-12 and 50), pattern counts at the tile boundaries (1, 63, 64, 65, 129),
all-gap columns, every schedule -- and the rejection of an empty alignment."""

import numpy as np
import pytest

from conftest import mcp, oracle, rel_err

pytestmark = pytest.mark.gpu

TREES = ["(A:0.1);", "(A:0.1,B:0.2);", "(A:0.1,B:0.2,C:0.3,D:0.05,E:0.4);",
         "((A:0.1,B:0.2):0.0,C:0.3);", "((A:1e-12,B:50.0):0.3,C:0.3);",
         "(((A:0.2,B:0.1):0.05,(C:0.3,D:0.01):0.2):0.1,E:0.6,F:0.02);"]


def _case(newick, S, C, R, seed, gaps):
    t = mcp.ParseNewick(newick)
    ta = mcp.tree_arrays(t)
    rng = np.random.default_rng(seed)
    x = np.zeros((S, C, ta.N), order="F")
    for leaf in ta.leaf_nums:
        x[rng.integers(0, S, C), np.arange(C), leaf - 1] = 1.0
        if gaps:
            x[:, rng.random(C) < 0.2, leaf - 1] = 1.0   # gaps / missing data: all states
    if S == 2:
        pi = np.array([0.3, 0.7])
        model = mcp.Restriction(pi)
    else:
        pi = rng.dirichlet(np.full(S, 5.0))
        model = mcp.GTR(pi, rng.exponential(1.0, S * (S - 1) // 2))
    rates = np.asarray(mcp.discrete_gamma_rates(0.5, 0.5, R)) if R > 1 else np.array([1.0])
    return t, ta, x, model, rates, pi


@pytest.mark.parametrize("newick", TREES)
@pytest.mark.parametrize("S,C,R,gaps", [(4, 1, 1, False), (4, 63, 3, True), (2, 64, 2, False),
                                         (4, 65, 4, True), (2, 129, 1, True)])
def test_edge_shapes_vs_oracle(newick, S, C, R, gaps):
    t, ta, x, (U, D, Uinv, mu), rates, pi = _case(newick, S, C, R, C + R, gaps)
    blv = mcp.get_branchlength_vector(t)
    ref_ll, ref_g = oracle.logl_grad(ta, x, U, D, Uinv, mu, rates, pi, blv)
    print(f"edge {newick} S={S} C={C} R={R} gaps={gaps}", flush=True)
    for sched in ("levels", "hybrid", "fused"):
        p = mcp.Plan(ta.parent, x, R)
        p.set_model(mcp.Model(U, D, Uinv, mu, rates, pi))
        p.set_schedule(sched)
        ll, g = p.logl_grad(blv)
        assert rel_err(ll, ref_ll) < 1e-12, (sched, ll, ref_ll)
        scale = max(1.0, float(np.max(np.abs(ref_g))))
        assert np.max(np.abs(g - ref_g)) <= 1e-10 * scale, (sched, g, ref_g)
        assert rel_err(p.logl(blv), ll) <= 1e-14
        p.close()


def test_empty_alignment_is_rejected():
    t, ta, x, (U, D, Uinv, mu), rates, pi = _case(TREES[1], 4, 4, 1, 0, False)
    with pytest.raises(mcp.McpError):
        mcp.Plan(ta.parent, x[:, :0, :], 1)
