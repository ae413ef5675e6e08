"""Branch-length priors of the tree samplers, with the reference's API
(src/Likelihood/Prior.jl, src/distributions/TreeDistribution.jl):

    exponentialBL(scale)               every branch Exponential(mean scale)
    CompoundDirichlet(alpha, a, beta, c)
        tree length T ~ Gamma(alpha, beta); branch-length proportions
        Dirichlet with a on leaf branches and a*c on internal branches
        (internal_logpdf, Prior.jl:1-37)

logpdf(d, tree) and gradlogpdf(d, tree) evaluate them on the host (O(N));
the device leapfrog evaluates the same densities in its second half
(mcp_hmc_set_prior_params; Batch.set_prior_compound_dirichlet).
"""

from __future__ import annotations

import math

import numpy as np

from .tree import Node, get_branchlength_vector, post_order

__all__ = ["exponentialBL", "CompoundDirichlet", "internal_external", "logpdf", "gradlogpdf"]


class exponentialBL:
    def __init__(self, scale: float, constraints=None):
        self.scale = float(scale)
        self.constraints = constraints


class CompoundDirichlet:
    def __init__(self, alpha: float, a: float, beta: float, c: float, constraints=None):
        self.alpha, self.a, self.beta, self.c = float(alpha), float(a), float(beta), float(c)
        self.constraints = constraints


def internal_external(tree: Node) -> np.ndarray:
    """1 for the branch above an internal node, 0 above a leaf, in the
    branch-length vector's order (node num 1..N-1)."""
    nodes = post_order(tree)
    N = len(nodes)
    out = np.zeros(N - 1, dtype=np.int64)
    for v in nodes:
        if v.num < N and v.children:
            out[v.num - 1] = 1
    return out


def _cd(d: CompoundDirichlet, blv, internal):
    blv = np.asarray(blv, dtype=np.float64)
    inl = np.asarray(internal) == 1
    lb = np.log(blv)
    T = float(blv.sum())
    nterm = float(np.count_nonzero(~inl))
    n_int = nterm - 3.0
    k = d.alpha - d.a * nterm - d.a * d.c * n_int
    lp = (d.alpha * math.log(d.beta) - math.lgamma(d.alpha) - T * d.beta
          - math.lgamma(d.a) - math.lgamma(d.c) + math.lgamma(d.a + d.c)
          + float(lb[~inl].sum()) * (d.a - 1.0) + float(lb[inl].sum()) * (d.a * d.c - 1.0)
          + k * math.log(T))
    g = -d.beta + np.where(inl, d.a * d.c - 1.0, d.a - 1.0) / blv + k / T
    return lp, g


def logpdf(d, tree: Node) -> float:
    if isinstance(d, CompoundDirichlet):
        return _cd(d, get_branchlength_vector(tree), internal_external(tree))[0]
    if isinstance(d, exponentialBL):
        b = get_branchlength_vector(tree)
        return float(np.sum(-math.log(d.scale) - b / d.scale))
    raise TypeError(f"no logpdf for {type(d).__name__}")


def gradlogpdf(d, tree: Node):
    if isinstance(d, CompoundDirichlet):
        return _cd(d, get_branchlength_vector(tree), internal_external(tree))
    if isinstance(d, exponentialBL):
        b = get_branchlength_vector(tree)
        return float(np.sum(-math.log(d.scale) - b / d.scale)), np.full(b.shape, -1.0 / d.scale)
    raise TypeError(f"no gradlogpdf for {type(d).__name__}")
