"""Seeded synthetic workloads: random trees, models and alignments simulated
down the tree.  These generate the benchmark configurations (20-taxon Yule
restriction-site tree; 100-taxon 4-state GTR with 4 gamma categories and
10,000 weighted patterns; ragged batches of 16..512-taxon trees; 1000-taxon
20-state trees) and the larger parity cases."""

from __future__ import annotations

import numpy as np

from .models import GTR, Restriction
from .rates import discrete_gamma_rates
from .tree import Node, number_nodes, post_order, tree_arrays

__all__ = ["yule_tree", "coalescent_tree", "caterpillar_tree", "simulate_alignment",
           "simulate_patterns",
           "random_gtr", "Workload", "make_workload", "transition_matrix"]

BASE_SEED = 20261202


def _finish(root: Node, rng, mean_bl: float) -> Node:
    for v in post_order(root):
        if v.mother is not None:
            v.inc_length = float(rng.exponential(mean_bl))
    leaves = [v for v in post_order(root) if not v.children]
    for k, v in enumerate(leaves):
        v.name = f"t{k}"
    number_nodes(root)
    return root


def yule_tree(n: int, rng, mean_bl: float = 0.1) -> Node:
    """Yule topology: start with two lineages below the root, split a
    uniformly chosen lineage until there are n."""
    root = Node()
    a, b = Node(), Node()
    root.add_child(a)
    root.add_child(b)
    tips = [a, b]
    while len(tips) < n:
        k = int(rng.integers(len(tips)))
        v = tips.pop(k)
        l, r = Node(), Node()
        v.add_child(l)
        v.add_child(r)
        tips += [l, r]
    return _finish(root, rng, mean_bl)


def coalescent_tree(n: int, rng, mean_bl: float = 0.05) -> Node:
    """Random rooted binary topology by random pairwise coalescence."""
    pool = [Node() for _ in range(n)]
    while len(pool) > 1:
        i, j = rng.choice(len(pool), size=2, replace=False)
        a, b = pool[i], pool[j]
        for k in sorted((i, j), reverse=True):
            pool.pop(k)
        v = Node()
        v.add_child(a)
        v.add_child(b)
        pool.append(v)
    return _finish(pool[0], rng, mean_bl)


def caterpillar_tree(n: int, rng, mean_bl: float = 0.1) -> Node:
    """Maximally unbalanced tree (n-1 post-order levels)."""
    v = Node()
    v.add_child(Node())
    v.add_child(Node())
    for _ in range(n - 2):
        w = Node()
        w.add_child(v)
        w.add_child(Node())
        v = w
    return _finish(v, rng, mean_bl)


def transition_matrix(U, D, Uinv, mu, rate, t) -> np.ndarray:
    return (U * np.exp(mu * t * D * rate)[None, :]) @ Uinv


def random_gtr(S: int, rng, exch_alpha: float = 1.0, freq_alpha: float = 5.0,
               exch: str = "dirichlet"):
    """Exchangeabilities ~ Dirichlet(alpha) (or iid Exp(1) with exch='exp') in
    setmatrix order, frequencies ~ Dirichlet(freq_alpha)."""
    m = S * (S - 1) // 2
    r = rng.dirichlet(np.full(m, exch_alpha)) if exch == "dirichlet" else rng.exponential(1.0, m)
    pi = rng.dirichlet(np.full(S, freq_alpha))
    return pi, r


def simulate_alignment(root: Node, model_fn, pi, subst_rates, rates, C: int, rng,
                       with_cats: bool = True) -> np.ndarray:
    """Simulate C sites root -> tips (root ~ pi, child ~ row of P); the rate
    category of a site is uniform over the R categories.  Returns one-hot tip
    partials (S, C, N), Fortran order."""
    U, D, Uinv, mu = model_fn(pi, subst_rates)
    S = len(pi)
    rates = np.atleast_1d(rates)
    R = rates.shape[0]
    cat = rng.integers(R, size=C) if (with_cats and R > 1) else np.zeros(C, dtype=np.int64)
    po = post_order(root)
    N = len(po)
    state = {}
    cdf_pi = np.cumsum(pi)
    state[root.num] = np.minimum(np.searchsorted(cdf_pi, rng.random(C) * cdf_pi[-1]), S - 1)
    for v in reversed(po):
        if v.mother is None:
            continue
        ps = state[v.mother.num]
        out = np.empty(C, dtype=np.int64)
        for r in range(R):
            sel = cat == r
            if not np.any(sel):
                continue
            P = transition_matrix(U, D, Uinv, mu, rates[r], v.inc_length)
            P = np.clip(P, 0.0, None)
            cdf = np.cumsum(P, axis=1)
            u = rng.random(int(sel.sum())) * cdf[ps[sel], -1]
            rows = cdf[ps[sel]]
            out[sel] = np.minimum((rows < u[:, None]).sum(axis=1), S - 1)
        state[v.num] = out
    x = np.zeros((S, C, N), order="F")
    for v in po:
        if not v.children:
            x[state[v.num], np.arange(C), v.num - 1] = 1.0
    return x


class Workload:
    """One (tree, data, model) problem: tree, x (S, C, N), weights, model
    tuple (U, D, Uinv, mu), rates, pi, and the substitution-model inputs."""

    def __init__(self, tree, x, weights, model_fn, pi, subst_rates, rates):
        self.tree = tree
        self.x = x
        self.weights = weights
        self.model_fn = model_fn
        self.pi = np.asarray(pi, dtype=np.float64)
        self.subst_rates = np.asarray(subst_rates, dtype=np.float64)
        self.rates = np.atleast_1d(np.asarray(rates, dtype=np.float64))
        self.U, self.D, self.Uinv, self.mu = model_fn(self.pi, self.subst_rates)
        self.ta = tree_arrays(tree)

    @property
    def S(self):
        return self.x.shape[0]

    @property
    def C(self):
        return self.x.shape[1]

    @property
    def N(self):
        return self.x.shape[2]


def simulate_patterns(root: Node, model_fn, pi, subst_rates, rates, target: int, rng):
    """Simulate sites in chunks until `target` distinct patterns exist; keep
    the first `target` distinct patterns (in order of first appearance) with
    weights = their counts among the simulated sites."""
    leaf = tree_arrays(root).leaf_nums - 1
    counts: dict[bytes, int] = {}
    cols: dict[bytes, np.ndarray] = {}
    order: list[bytes] = []
    x_all = None
    while len(order) < target:
        x = simulate_alignment(root, model_fn, pi, subst_rates, rates, max(target, 1024), rng)
        if x_all is None:
            x_all = x
        codes = np.argmax(x[:, :, leaf], axis=0)          # (C, L) leaf states
        for c in range(codes.shape[0]):
            k = codes[c].astype(np.int8).tobytes()
            if k in counts:
                counts[k] += 1
            elif len(order) < target:
                counts[k] = 1
                cols[k] = x[:, c, :]
                order.append(k)
    S, _, N = x_all.shape
    xu = np.zeros((S, target, N), order="F")
    for j, k in enumerate(order):
        xu[:, j, :] = cols[k]
    w = np.array([counts[k] for k in order], dtype=np.float64)
    return xu, w


def make_workload(name: str, seed: int | None = None, **kw) -> Workload:
    """Named workloads.

    restriction20  20-taxon Yule, t ~ Exp(0.1), restriction model pi=(.5,.5),
                   1000 simulated sites, 1 category
    gtr100         100 taxa random topology, t ~ Exp(0.05), 4-state GTR
                   (exch ~ Dir(1), pi ~ Dir(5)), 4 gamma categories (alpha .5),
                   10,000 distinct simulated patterns with count weights
    gtr200         200 taxa, GTR, 5000 patterns, 4 categories
    aa1000         1000 taxa, 20 states, exch iid Exp(1), pi ~ Dir(10),
                   4 categories, `C` patterns (default 100,000)
    custom         n, S, C, R from keywords
    """
    rng = np.random.default_rng(BASE_SEED if seed is None else seed)
    if name == "restriction20":
        t = yule_tree(kw.get("n", 20), rng, 0.1)
        pi = np.array([0.5, 0.5])
        x = simulate_alignment(t, Restriction, pi, np.zeros(1), [1.0], kw.get("C", 1000), rng)
        return Workload(t, x, None, Restriction, pi, np.zeros(1), [1.0])
    if name in ("gtr100", "gtr200"):
        n = kw.get("n", 100 if name == "gtr100" else 200)
        C = kw.get("C", 10_000 if name == "gtr100" else 5000)
        t = coalescent_tree(n, rng, 0.05)
        pi, r = random_gtr(4, rng, 1.0, 5.0)
        rates = discrete_gamma_rates(0.5, 0.5, 4)
        xu, w = simulate_patterns(t, GTR, pi, r, rates, C, rng)
        return Workload(t, xu, w, GTR, pi, r, rates)
    if name == "aa1000":
        n = kw.get("n", 1000)
        C = kw.get("C", 100_000)
        t = coalescent_tree(n, rng, 0.05)
        pi, r = random_gtr(20, rng, 1.0, 10.0, exch="exp")
        rates = discrete_gamma_rates(0.5, 0.5, 4)
        x = simulate_alignment(t, GTR, pi, r, rates, C, rng)
        return Workload(t, x, None, GTR, pi, r, rates)
    if name == "custom":
        n, S, C, R = kw["n"], kw["S"], kw["C"], kw.get("R", 1)
        topo = kw.get("topology", "coalescent")
        t = {"coalescent": coalescent_tree, "yule": yule_tree,
             "caterpillar": caterpillar_tree}[topo](n, rng, kw.get("mean_bl", 0.1))
        if S == 2:
            pi = rng.dirichlet([5.0, 5.0])
            model_fn, r = Restriction, np.zeros(1)
        else:
            pi, r = random_gtr(S, rng)
            model_fn = GTR
        rates = discrete_gamma_rates(0.5, 0.5, R) if R > 1 else np.array([1.0])
        x = simulate_alignment(t, model_fn, pi, r, rates, C, rng)
        return Workload(t, x, None, model_fn, pi, r, rates)
    raise ValueError(f"unknown workload {name!r}")
