"""Substitution-model eigensystems (host side).

Each model returns (U, D, Uinv, mu) with Q = U diag(D) Uinv and the
expected-changes normaliser mu, exactly as the reference's models do
(`src/Likelihood/SubstitutionModels.jl:13-113`); the likelihood path then
builds P(t) = U diag(exp(mu rho t D)) Uinv on the device.  Like the reference
(Julia `LinearAlgebra.eigen` + `inv`), the decomposition is LAPACK's general
eigen-solver followed by a matrix inverse; eigenvalues are sorted ascending
(Julia's `eigen` sorts real spectra the same way).  Eigenvector sign and scale
are LAPACK-dependent and do not change P.
"""

from __future__ import annotations

import math

import numpy as np

__all__ = ["Restriction", "JC", "GTR", "freeK", "setmatrix", "ComplexEigensystemError"]


class ComplexEigensystemError(ValueError):
    """Q has a complex spectrum; the reference fails on the Float64 store of
    such a P (`VectorizedFunctions.jl:155-168`), so this path rejects it."""


def _recip(x) -> float:
    """1/x with IEEE semantics (1/0 = Inf), as Julia computes it."""
    with np.errstate(divide="ignore"):
        return float(np.float64(1.0) / np.float64(x))


def _eigen(Q: np.ndarray):
    D, U = np.linalg.eig(Q)
    if np.iscomplexobj(D):
        scale = max(1.0, float(np.max(np.abs(D))))
        if np.max(np.abs(D.imag)) > 1e-12 * scale or np.max(np.abs(U.imag)) > 1e-12:
            raise ComplexEigensystemError("substitution matrix has a complex eigensystem")
        D, U = D.real, U.real
    order = np.argsort(D, kind="stable")
    D = np.ascontiguousarray(D[order])
    U = np.asfortranarray(U[:, order])
    Uinv = np.asfortranarray(np.linalg.inv(U))
    return U, D, Uinv


def Restriction(base_freq, substitution_rates=None):
    """Binary restriction-site model (`SubstitutionModels.jl:13-24`).
    `substitution_rates` is ignored, as in the reference."""
    f = np.asarray(base_freq, dtype=np.float64)
    if f.shape != (2,):
        raise AssertionError("Restriction needs exactly two base frequencies")
    D = np.array([-1.0, 0.0])
    U = np.array([[f[1], f[1]], [-1.0 + f[1], 0.0 + f[1]]], order="F")
    Uinv = np.asfortranarray(np.linalg.inv(U))
    mu = _recip(2.0 * float(np.prod(f)))
    return U, D, Uinv, mu


def JC(base_freq, substitution_rates=None):
    """Jukes-Cantor over len(base_freq) states (`SubstitutionModels.jl:36-50`).
    mu = 1 / ((S-1)/S)."""
    S = len(base_freq)
    off = 1.0 / S
    diag = off * (S - 1)
    Q = np.full((S, S), off)
    Q[np.diag_indices(S)] = -diag
    U, D, Uinv = _eigen(Q)
    return U, D, Uinv, _recip(diag)


def setmatrix(vec_vals) -> np.ndarray:
    """Symmetric matrix from its strict upper triangle filled column by column:
    (1,2), (1,3), (2,3), (1,4), ... (`SubstitutionModels.jl:106-113`)."""
    v = np.asarray(vec_vals, dtype=np.float64)
    n = v.shape[0]
    s = int(round((math.sqrt(8 * n + 1) + 1) / 2))
    if s * (s - 1) // 2 != n:
        raise ValueError("setmatrix: length of vector is not triangular")
    Q = np.zeros((s, s))
    k = 0
    for j in range(s):
        for i in range(j):
            Q[i, j] = v[k]
            k += 1
    return Q + Q.T


def GTR(base_freq, substitution_rates):
    """General time-reversible model (`SubstitutionModels.jl:62-74`):
    Q[i,j] = R[i,j] * pi[j], Q[i,i] = -row sum, mu = 1 / sum of the row sums
    (not pi-weighted)."""
    pi = np.asarray(base_freq, dtype=np.float64)
    S = pi.shape[0]
    Q = setmatrix(substitution_rates)
    if Q.shape != (S, S):
        raise ValueError("GTR: number of exchangeabilities does not match the number of states")
    Q = Q * pi[None, :]
    dia = Q.sum(axis=1)
    Q[np.diag_indices(S)] = -dia
    U, D, Uinv = _eigen(Q)
    return U, D, Uinv, _recip(dia.sum())


def freeK(base_freq, substitution_rates):
    """Free (non-reversible) rate matrix (`SubstitutionModels.jl:83-102`):
    off-diagonal Q[j,i] filled from the rate vector with i outer, j inner."""
    r = np.asarray(substitution_rates, dtype=np.float64).ravel()
    S = int(math.ceil(math.sqrt(r.shape[0])))
    Q = np.zeros((S, S))
    k = 0
    for i in range(S):
        for j in range(S):
            if i != j:
                Q[j, i] = r[k]
                k += 1
    dia = Q.sum(axis=1)
    Q[np.diag_indices(S)] = -dia
    U, D, Uinv = _eigen(Q)
    return U, D, Uinv, _recip(dia.sum())


def stationary_freeK(substitution_rates) -> np.ndarray:
    """Equilibrium frequencies used by the freeK PhyloDist constructor
    (`src/distributions/Phylodist.jl:89-100`): first row of U diag(e_S) Uinv."""
    U, D, Uinv, _ = freeK(np.zeros(0), substitution_rates)
    E = np.zeros_like(D)
    E[-1] = 1.0
    return np.real((U @ np.diag(E) @ Uinv)[0, :])
