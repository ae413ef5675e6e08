"""Independent brute-force likelihood for small trees (test helper).

Sums pi_root * prod_edges P over every assignment of states to the internal
nodes -- no pruning, no rescaling, no eigen-decomposition (P from
scipy.linalg.expm of the rate matrix built directly from the model's
definition).  Per (site, category) the likelihood is that sum; the total is
sum_{c,r} w_c log L_{c,r}, the reference's category semantics
(VectorizedFunctions.jl:37-40,76-87).
"""

import itertools

import numpy as np
from scipy.linalg import expm


def rate_matrix(kind, pi, subst_rates):
    """Q (unnormalised) and mu exactly as defined by the reference models
    (SubstitutionModels.jl:13-102), built without eigen-decomposition."""
    pi = np.asarray(pi, dtype=np.float64)
    S = pi.shape[0]
    if kind == "JC":
        Q = np.full((S, S), 1.0 / S)
        np.fill_diagonal(Q, -(S - 1) / S)
        return Q, S / (S - 1)
    if kind == "Restriction":
        f1, f2 = pi
        Q = np.array([[-f2, f2], [f1, -f1]])  # eigenvalues -1, 0 with pi stationary
        return Q, 1.0 / (2.0 * f1 * f2)
    if kind == "GTR":
        r = np.asarray(subst_rates, dtype=np.float64)
        Rm = np.zeros((S, S))
        k = 0
        for j in range(S):
            for i in range(j):
                Rm[i, j] = Rm[j, i] = r[k]
                k += 1
        Q = Rm * pi[None, :]
        d = Q.sum(axis=1)
        np.fill_diagonal(Q, -d)
        return Q, 1.0 / d.sum()
    raise ValueError(kind)


def brute_logl(ta, x, Q, mu, rates, pi, blv, weights=None, grad=False):
    """ll (and d ll / d blv by brute force differentiation of each edge)."""
    S, C, N = x.shape
    rates = np.atleast_1d(rates)
    w = np.ones(C) if weights is None else np.asarray(weights)
    nch = np.diff(ta.child_ptr)
    internal = [v for v in range(1, N + 1) if nch[v - 1] > 0]
    root = N
    edges = [v for v in range(1, N) if True]
    ll = 0.0
    g = np.zeros(N - 1)
    for rho in rates:
        P = {v: expm(mu * rho * blv[v - 1] * Q) for v in edges}
        dP = {v: mu * rho * Q @ P[v] for v in edges}
        L = np.zeros(C)
        dL = np.zeros((N - 1, C))
        for assign in itertools.product(range(S), repeat=len(internal)):
            st = dict(zip(internal, assign))
            factors = {}
            for v in edges:
                p = ta.parent[v - 1]
                if nch[v - 1] > 0:
                    factors[v] = np.full(C, P[v][st[p], st[v]])
                else:
                    factors[v] = P[v][st[p], :] @ x[:, :, v - 1]
            term = np.full(C, pi[st[root]])
            for v in edges:
                term = term * factors[v]
            L += term
            if grad:
                for v in edges:
                    p = ta.parent[v - 1]
                    if nch[v - 1] > 0:
                        dfac = np.full(C, dP[v][st[p], st[v]])
                    else:
                        dfac = dP[v][st[p], :] @ x[:, :, v - 1]
                    rest = np.full(C, pi[st[root]])
                    for u in edges:
                        if u != v:
                            rest = rest * factors[u]
                    dL[v - 1] += rest * dfac
        ll += float(np.sum(w * np.log(L)))
        if grad:
            g += (dL / L[None, :]) @ w
    return (ll, g) if grad else ll
