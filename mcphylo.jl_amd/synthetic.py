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


# ---------------------------------------------------------------------------
# ragged batches (many trees x chains), generated as flat arrays
# ---------------------------------------------------------------------------

def coalescent_parent(n: int, rng):
    """Random rooted binary topology by pairwise coalescence as a parent
    array: leaves 1..n, internal nodes n+1..2n-1 in merge order (children
    before parents), root 2n-1."""
    N = 2 * n - 1
    parent = np.zeros(N, dtype=np.int32)
    pool = list(range(1, n + 1))
    nxt = n + 1
    while len(pool) > 1:
        i, j = rng.choice(len(pool), size=2, replace=False)
        a, b = pool[i], pool[j]
        for k in sorted((i, j), reverse=True):
            pool.pop(k)
        parent[a - 1] = nxt
        parent[b - 1] = nxt
        pool.append(nxt)
        nxt += 1
    return parent


def simulate_codes(parent, blv, U, D, Uinv, mu, rates, pi, C: int, rng) -> np.ndarray:
    """Leaf state codes (n_leaves x C, uint8, leaves in increasing num order)
    simulated down the tree; the category of a site is uniform over R."""
    N = parent.shape[0]
    S = len(pi)
    rates = np.atleast_1d(rates)
    R = rates.shape[0]
    cat = rng.integers(R, size=C)
    # parents before children: order by depth
    depth = np.zeros(N, dtype=np.int64)
    order = [N - 1]
    kids = [[] for _ in range(N)]
    for v in range(N - 1):
        kids[parent[v] - 1].append(v)
    k = 0
    while k < len(order):
        order.extend(kids[order[k]])
        k += 1
    state = np.zeros((N, C), dtype=np.uint8 if S <= 255 else np.int64)
    cdf_pi = np.cumsum(pi)
    state[N - 1] = np.minimum(np.searchsorted(cdf_pi, rng.random(C) * cdf_pi[-1]), S - 1)
    for v in order[1:]:
        E = np.exp(mu * blv[v] * D[None, :] * rates[:, None])         # (R, S)
        P = np.clip((U[None, :, :] * E[:, None, :]) @ Uinv, 0, None)      # (R, S, S)
        cdf = np.cumsum(P, axis=2)
        rows = cdf[cat, state[parent[v] - 1]]                            # (C, S)
        u = rng.random(C) * rows[:, -1]
        state[v] = np.minimum((rows < u[:, None]).sum(axis=1), S - 1)
    leaves = np.array([v for v in range(N) if not kids[v]])
    return state[leaves].astype(np.uint8), leaves + 1


def _batched_tree(i: int, seed: int, n_range, c_range, R: int, chains: int):
    """Tree i of a ragged batch (see make_batched_workload)."""
    rates = discrete_gamma_rates(0.5, 0.5, R) if R > 1 else np.array([1.0])
    rng = np.random.default_rng(seed + i)
    n = int(rng.integers(n_range[0], n_range[1] + 1))
    C = int(round(np.exp(rng.uniform(np.log(c_range[0]), np.log(c_range[1])))))
    if rng.random() < 0.5:
        pi = rng.dirichlet([5.0, 5.0])
        U, D, Ui, mu = Restriction(pi)
    else:
        pi, r = random_gtr(4, rng, 1.0, 5.0)
        U, D, Ui, mu = GTR(pi, r)
    S = len(pi)
    parent = coalescent_parent(n, rng)
    blv = rng.exponential(0.05, 2 * n - 2)
    codes, leaf_nums = simulate_codes(parent, blv, U, D, Ui, mu, rates, pi, C, rng)
    table = np.vstack([np.eye(S), np.ones((1, S))])
    ds = {"leaf_nums": leaf_nums, "tip_codes": codes, "code_table": table, "weights": None}
    model = (U, D, Ui, mu, rates, pi)
    blvs = [blv * np.exp(rng.normal(0.0, 0.1, blv.shape)) for _ in range(chains)]
    return ds, parent, model, blvs


class DenseWorkload:
    """A large many-state workload held as leaf state codes (the one-hot tip
    partials of 1000 taxa x 100,000 sites x 20 states would take 32 GB)."""

    def __init__(self, parent, blv, codes, leaf_nums, table, model):
        self.parent, self.blv, self.codes, self.leaf_nums = parent, blv, codes, leaf_nums
        self.table = table
        self.U, self.D, self.Uinv, self.mu, self.rates, self.pi = model
        self.S = table.shape[1]
        self.C = codes.shape[1]
        self.N = parent.shape[0]
        self.R = len(self.rates)

    def dataset(self, cols=None):
        """Batch dataset dict over all sites or the site slice `cols`."""
        codes = self.codes if cols is None else np.ascontiguousarray(self.codes[:, cols])
        return {"leaf_nums": self.leaf_nums, "tip_codes": codes, "code_table": self.table,
                "weights": None}

    def tip_partials(self, cols):
        """One-hot (S, C, N) tip partials of the site slice (for the oracle)."""
        codes = self.codes[:, cols]
        x = np.zeros((self.S, codes.shape[1], self.N), order="F")
        for k, v in enumerate(self.leaf_nums):
            x[:, :, v - 1] = self.table[codes[k]].T
        return x


def make_dense_workload(n: int = 1000, C: int = 100_000, S: int = 20, R: int = 4,
                        seed: int = BASE_SEED) -> DenseWorkload:
    """n taxa (random pairwise coalescence), t ~ Exp(0.05), S-state model with
    exchangeabilities iid Exp(1) and pi ~ Dir(10), R gamma categories (alpha
    0.5), C sites simulated down the tree as leaf state codes."""
    rng = np.random.default_rng(seed)
    parent = coalescent_parent(n, rng)
    blv = rng.exponential(0.05, 2 * n - 2)
    pi, r = random_gtr(S, rng, 1.0, 10.0, exch="exp")
    U, D, Ui, mu = GTR(pi, r)
    rates = discrete_gamma_rates(0.5, 0.5, R) if R > 1 else np.array([1.0])
    codes, leaf_nums = simulate_codes(parent, blv, U, D, Ui, mu, rates, pi, C, rng)
    table = np.vstack([np.eye(S), np.ones((1, S))])
    return DenseWorkload(parent, blv, codes, leaf_nums, table, (U, D, Ui, mu, rates, pi))


def make_batched_workload(n_trees: int = 256, chains: int = 4, seed: int = BASE_SEED,
                          n_range=(16, 512), c_range=(500, 20000), R: int = 4, pool=None):
    """Uneven batch: tree i (seed+i) has n ~ U{n_range}, C = round(exp(U[ln c_range]))
    patterns simulated down the tree, a restriction-site (pi ~ Dir(5,5)) or
    4-state GTR model (50/50), R gamma categories (alpha 0.5), t ~ Exp(0.05);
    `chains` items per tree whose branch lengths are the tree's times
    exp(N(0, 0.1^2)).  `pool` (an Executor) generates trees in parallel.
    Returns (datasets, items, models, blv_concat) for engine.Batch."""
    args = [(i, seed, n_range, c_range, R, chains) for i in range(n_trees)]
    if pool is None:
        res = [_batched_tree(*a) for a in args]
    else:
        res = list(pool.map(_batched_tree, *zip(*args)))
    datasets, items, models, blvs = [], [], [], []
    for i, (ds, parent, model, bl) in enumerate(res):
        datasets.append(ds)
        for b in bl:
            items.append({"dataset": i, "parent": parent, "n_rates": R})
            models.append(model)
            blvs.append(b)
    return datasets, items, models, np.concatenate(blvs)
