"""PhyloDist / MultiplePhyloDist: the reference's distribution API, backed by
the HIP engine.

Mirrors `src/distributions/Phylodist.jl`:
  PhyloDist struct and constructors      :8-100
  size(d) == (nbase, 1, nnodes)          :105
  logpdf(d::PhyloDist, x)                :107-122
  gradlogpdf(d::PhyloDist, x)            :124-138
  MultiplePhyloDist + constructors       :143-272
  logpdf(d::MultiplePhyloDist, x)        :281-288
  __logpdf(d::MultiplePhyloDist, x)      :290-297

The reference rebuilds a PhyloDist on every likelihood call
(`src/model/dependent.jl:467-471`), so device plans are cached by the
identity of the data array `x` plus the tree topology, not by the
PhyloDist instance.  Like the reference, `x` must not be mutated in place
between calls (use `clear_plan_cache()` after doing so).  Per call only the
branch lengths (and the model, when it changed) travel to the device.
"""

from __future__ import annotations

import weakref

import numpy as np

from . import _lib
from .engine import Batch, Model, Plan
from .models import (GTR, JC, Restriction, freeK, gtr_gradient, rate_gradient,
                     restriction_gradient, stationary_freeK)
from .tree import Node, get_branchlength_vector, post_order, tree_arrays

__all__ = ["PhyloDist", "MultiplePhyloDist", "logpdf", "gradlogpdf", "gradlogpdf_params",
           "__logpdf", "logpdf_multi", "clear_plan_cache", "size", "minimum", "maximum"]


class PhyloDist:
    """Distribution whose likelihood is Felsenstein's pruning algorithm.

    PhyloDist(tree, base_freq, substitution_rates, rates, substitution_model)
    PhyloDist(tree, substitution_rates, rates, freeK)   # pi from freeK's equilibrium

    Build extensions (keyword-only): `pattern_weights` [C] multiply every
    per-pattern term; `compress` merges identical columns into weighted
    patterns on the device plan (default True; sums change only in summation
    order); `device` selects the GPU.
    """

    def __init__(self, tree: Node, *args, pattern_weights=None, compress: bool = True,
                 device: int = 0):
        if len(args) == 3 and args[2] is freeK:
            substitution_rates, rates, model = args
            base_freq = stationary_freeK(np.asarray(substitution_rates, dtype=np.float64))
        elif len(args) == 4:
            base_freq, substitution_rates, rates, model = args
        else:
            raise TypeError("PhyloDist(tree, base_freq, substitution_rates, rates, model) or "
                            "PhyloDist(tree, substitution_rates, rates, freeK)")
        self.tree = tree
        self.base_freq = np.array(np.atleast_1d(base_freq), dtype=np.float64)
        self.substitution_rates = np.array(np.atleast_1d(substitution_rates), dtype=np.float64)
        self.rates = np.array(np.atleast_1d(rates), dtype=np.float64)
        self.substitution_model = model
        self.nbase = int(self.base_freq.shape[0])
        self.nnodes = len(post_order(tree))
        self.pattern_weights = None if pattern_weights is None else np.asarray(
            pattern_weights, dtype=np.float64)
        self.compress = bool(compress)
        self.device = int(device)

    def __eq__(self, other):
        if not isinstance(other, PhyloDist):
            return NotImplemented
        return (self.tree is other.tree and np.array_equal(self.base_freq, other.base_freq) and
                np.array_equal(self.substitution_rates, other.substitution_rates) and
                np.array_equal(self.rates, other.rates) and
                self.substitution_model is other.substitution_model)

    def model(self) -> Model:
        U, D, Uinv, mu = self.substitution_model(self.base_freq, self.substitution_rates)
        return Model(U, D, Uinv, mu, self.rates, self.base_freq)


def size(d):
    if isinstance(d, MultiplePhyloDist):
        return (d.DistCollector[0].base_freq.shape[0], 1, int(max(d.size_array)),
                len(d.size_array))
    return (d.nbase, 1, d.nnodes)


def minimum(d):
    return -np.inf


def maximum(d):
    return np.inf


# ---------------------------------------------------------------------------
# plan cache: data identity + topology -> device plan
# ---------------------------------------------------------------------------

class _CacheEntry:
    __slots__ = ("ref", "plans")

    def __init__(self, ref):
        self.ref = ref
        self.plans = {}


_cache: dict[int, _CacheEntry] = {}


def clear_plan_cache() -> None:
    for e in list(_cache.values()):
        for p in e.plans.values():
            p.close()
    _cache.clear()


def _entry_for(x: np.ndarray) -> _CacheEntry:
    k = id(x)
    e = _cache.get(k)
    if e is not None and e.ref() is x:
        return e
    if e is not None:
        for p in e.plans.values():
            p.close()
    try:
        ref = weakref.ref(x, lambda _r, k=k: _evict(k))
    except TypeError:
        ref = (lambda x=x: x)
    e = _CacheEntry(ref)
    _cache[k] = e
    return e


def _evict(k: int) -> None:
    e = _cache.pop(k, None)
    if e is not None:
        for p in e.plans.values():
            try:
                p.close()
            except Exception:
                pass


def _plan_for(d: PhyloDist, x: np.ndarray):
    if x.ndim != 3:
        raise _lib.McpDimensionMismatch(_lib.MCP_ERR_DIMENSION, "x must be (S, C, N)")
    S, C, N = x.shape
    if S != d.nbase or N != d.nnodes:
        raise _lib.McpDimensionMismatch(_lib.MCP_ERR_DIMENSION,
                                        f"x has shape {x.shape}, expected ({d.nbase}, C, {d.nnodes})")
    ta = tree_arrays(d.tree)
    e = _entry_for(x)
    leaves = tuple(ta.leaf_nums.tolist())
    wkey = None if d.pattern_weights is None else d.pattern_weights.tobytes()
    key = (leaves, d.rates.shape[0], d.device, d.compress, wkey)
    plan = e.plans.get(key)
    if plan is None:
        plan = Plan(ta.parent, x, d.rates.shape[0], weights=d.pattern_weights, device=d.device,
                    compress=d.compress)
        e.plans[key] = plan
    elif not np.array_equal(plan.parent, ta.parent):
        plan.set_topology(ta.parent)
    plan.set_model(d.model())
    return plan


def logpdf(d, x) -> float:
    """logpdf(d::PhyloDist, x) -> Float64 (`Phylodist.jl:107-122`), or the sum
    over trees for a MultiplePhyloDist (`Phylodist.jl:281-288`)."""
    if isinstance(d, MultiplePhyloDist):
        return logpdf_multi(d, x)
    plan = _plan_for(d, x)
    return plan.logl(get_branchlength_vector(d.tree))


def gradlogpdf(d: PhyloDist, x):
    """gradlogpdf(d::PhyloDist, x) -> (ll, grad[N-1] by num)
    (`Phylodist.jl:124-138`)."""
    plan = _plan_for(d, x)
    return plan.logl_grad(get_branchlength_vector(d.tree))


def gradlogpdf_params(d: PhyloDist, x):
    """(ll, {name: gradient}) with respect to the branch lengths and the
    model's own parameters: 'branch_lengths' [N-1] by num, 'rates' [R]
    (category rates), 'base_freq' [S] and 'substitution_rates' (GTR
    exchangeabilities in setmatrix order; zeros for models that ignore them).
    GTR, Restriction and JC with S <= 8.  No reference counterpart: the
    reference differentiates the branch lengths only (`Phylodist.jl:124-138`)."""
    if d.substitution_model not in (GTR, Restriction, JC):
        raise _lib.McpError(_lib.MCP_ERR_UNSUPPORTED,
                            "parameter gradients cover the GTR, Restriction and JC models")
    plan = _plan_for(d, x)
    blv = get_branchlength_vector(d.tree)
    ll, g, dP, dpi = plan.logl_grad_model(blv)
    U, D, Uinv, mu = d.substitution_model(d.base_freq, d.substitution_rates)
    out = {"branch_lengths": g, "rates": rate_gradient(U, D, Uinv, mu, blv, d.rates, dP),
           "substitution_rates": np.zeros_like(d.substitution_rates)}
    if d.substitution_model is GTR:
        out["substitution_rates"], out["base_freq"] = gtr_gradient(
            d.base_freq, d.substitution_rates, U, D, Uinv, mu, blv, d.rates, dP, dpi)
    elif d.substitution_model is Restriction:
        out["base_freq"] = restriction_gradient(d.base_freq, U, D, Uinv, mu, blv, d.rates, dP,
                                                dpi)
    else:
        out["base_freq"] = np.asarray(dpi, dtype=np.float64).copy()  # JC: Q ignores pi
    return ll, out


# ---------------------------------------------------------------------------
# MultiplePhyloDist
# ---------------------------------------------------------------------------

class MultiplePhyloDist:
    """Collection of PhyloDists over several trees sharing S and C
    (`Phylodist.jl:143-272`).  Evaluated as one batch on the device.

    MultiplePhyloDist(trees, base_freq, substitution_rates, rates, model)
      base_freq / substitution_rates / rates: 1-D (shared) or 2-D with one
      column per tree; DimensionMismatch otherwise.
    MultiplePhyloDist(trees, substitution_rates, rates, freeK)
    """

    def __init__(self, tree_array, *args, device: int = 0):
        trees = list(tree_array)
        n_t = len(trees)

        def cols(a, name):
            a = np.asarray(a, dtype=np.float64)
            if a.ndim == 1:
                return np.repeat(a[:, None], n_t, axis=1)
            if a.shape[1] == n_t:
                return a
            if a.shape[1] == 1:
                return np.repeat(a, n_t, axis=1)
            raise _lib.McpDimensionMismatch(_lib.MCP_ERR_DIMENSION,
                                            f"Size of {name} and tree_array are incompatible")

        if len(args) == 3 and args[2] is freeK:
            sr, rates, model = args
            sr = cols(sr, "substitution_rates")
            bf = np.stack([stationary_freeK(sr[:, i]) for i in range(n_t)], axis=1)
            rates = np.asarray(rates, dtype=np.float64)
            rates = rates[:, None].repeat(n_t, axis=1) if rates.ndim == 1 else rates
        elif len(args) == 4:
            bf, sr, rates, model = args
            bf = cols(bf, "base_freq")
            sr = cols(sr, "substitution_rates")
            rates = cols(rates, "rates")
        else:
            raise TypeError("MultiplePhyloDist(trees, base_freq, substitution_rates, rates, model)")
        self.DistCollector = [PhyloDist(t, bf[:, i], sr[:, i], rates[:, min(i, rates.shape[1] - 1)],
                                        model, device=device) for i, t in enumerate(trees)]
        self.size_array = np.array([len(post_order(t)) for t in trees], dtype=np.int64)
        self.device = device
        self._batch = None
        self._batch_key = None


def _batch_for(d: MultiplePhyloDist, x: np.ndarray):
    if x.ndim != 4 or x.shape[3] != len(d.size_array):
        raise _lib.McpDimensionMismatch(_lib.MCP_ERR_DIMENSION, "x must be (S, C, maxN, ntrees)")
    tas = [tree_arrays(pd.tree) for pd in d.DistCollector]
    key = (id(x), tuple(ta.topology_key() for ta in tas))
    if d._batch is None or d._batch_key != key:
        datasets, items = [], []
        for i, (pd, ta) in enumerate(zip(d.DistCollector, tas)):
            s = int(d.size_array[i])
            xt = x[:, :, :s, i]
            datasets.append({"x_leaves": xt[:, :, ta.leaf_nums - 1], "leaf_nums": ta.leaf_nums,
                             "weights": None, "compress": True})
            items.append({"dataset": i, "parent": ta.parent, "n_rates": pd.rates.shape[0]})
        if d._batch is not None:
            d._batch.close()
        d._batch = Batch(datasets, items, device=d.device)
        d._batch_key = key
    for i, pd in enumerate(d.DistCollector):
        d._batch.set_model(i, pd.model())
    blv = np.concatenate([get_branchlength_vector(pd.tree) for pd in d.DistCollector])
    return d._batch, blv


def logpdf_multi(d: MultiplePhyloDist, x) -> float:
    b, blv = _batch_for(d, x)
    return float(np.sum(b.logl(blv)))


def __logpdf(d: MultiplePhyloDist, x):
    """Per-tree (ll, grad) tuples (`Phylodist.jl:290-297`)."""
    b, blv = _batch_for(d, x)
    ll, g = b.logl_grad(blv)
    return [(float(ll[i]), gi) for i, gi in enumerate(b.split(g))]
