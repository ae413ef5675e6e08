"""GPU: the dense 20-state workload at its full size -- 1000 taxa, 100,000
sites as leaf state codes, 4 gamma categories, FP64-MFMA level kernels.
Parity with the CPU oracle on a site slice, and at full size the
size-independent property that logL and every branch gradient are sums over
disjoint site chunks (exact up to summation order)."""

import time

import numpy as np
import pytest

from conftest import mcp, oracle, rel_err

pytestmark = pytest.mark.gpu

LL_TOL, GRAD_TOL = 1e-10, 1e-8


def grad_err(g, ref):
    """Max abs difference relative to the largest reference entry (floor 1e-3)."""
    return float(np.max(np.abs(g - ref)) / max(np.max(np.abs(ref)), 1e-3))


@pytest.fixture(scope="module")
def dense():
    t0 = time.perf_counter()
    w = mcp.synthetic.make_dense_workload()
    print(f"dense workload generated in {time.perf_counter() - t0:.1f}s", flush=True)
    return w


def _batch(w, cols=None, schedule="levels"):
    b = mcp.Batch([w.dataset(cols)], [{"dataset": 0, "parent": w.parent, "n_rates": w.R}])
    b.set_model(0, mcp.Model(w.U, w.D, w.Uinv, w.mu, w.rates, w.pi))
    b.set_schedule(schedule)
    return b


@pytest.mark.parametrize("schedule", ["levels", "fused"])
def test_dense20_site_slice_vs_oracle(dense, schedule):
    w = dense
    cols = slice(1000, 1300)
    x = w.tip_partials(cols)
    ta = mcp.tree_arrays(mcp.tree_from_parent(w.parent, w.blv))
    llo, go = oracle.logl_grad(ta, x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, w.blv)
    b = _batch(w, cols, schedule)
    ll, g = b.logl_grad(w.blv)
    assert rel_err(ll[0], llo) < LL_TOL
    assert grad_err(g, go) < GRAD_TOL
    b.close()


def test_dense20_full_size_is_sum_of_site_chunks(dense):
    w = dense
    assert (w.N, w.S, w.C, w.R) == (1999, 20, 100_000, 4)
    t0 = time.perf_counter()
    b = _batch(w)
    ll, g = b.logl_grad(w.blv)
    t1 = time.perf_counter()
    ll2 = b.logl(w.blv)
    print(f"full size: plan+first eval {t1 - t0:.2f}s, ll {ll[0]:.6f}", flush=True)
    assert ll2[0] == ll[0]
    assert np.all(np.isfinite(g)) and np.isfinite(ll[0])
    b.close()
    lsum, gsum = 0.0, np.zeros_like(g)
    for a in range(0, w.C, 25_000):
        bc = _batch(w, slice(a, a + 25_000))
        lc, gc = bc.logl_grad(w.blv)
        bc.close()
        lsum += lc[0]
        gsum += gc
    assert rel_err(ll[0], lsum) < 1e-12
    assert grad_err(g, gsum) < 1e-10
