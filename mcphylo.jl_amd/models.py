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

__all__ = ["Restriction", "JC", "GTR", "freeK", "setmatrix", "ComplexEigensystemError",
           "dlogl_dM", "gtr_gradient", "restriction_gradient", "rate_gradient"]


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


# ---------------------------------------------------------------------------
# Gradient with respect to the substitution model's parameters
# ---------------------------------------------------------------------------
# The device returns d logL / d P_{v,r} (every branch v and category r) and
# d logL / d pi through the root term.  P_{v,r} = exp(tau M) with tau = rho_r
# t_v and M = mu Q = U diag(mu D) Uinv, so by the Daleckii-Krein formula
#   dP = U [(Uinv dM U) o Phi] Uinv,  Phi_ij = tau (e^{a_i} - e^{a_j}) / (a_i - a_j),
# a = tau mu D (Phi_ii = tau e^{a_i}), and
#   d logL / dM = Uinv^T [sum_{v,r} (U^T A_{v,r} Uinv^T) o Phi_{v,r}] U^T.
# The model's parameters then enter through M = mu(theta) Q(theta).


def _exp_divided_difference(a: np.ndarray) -> np.ndarray:
    """(e^{a_i} - e^{a_j}) / (a_i - a_j) over the last axis pairs (e^{a_i}
    on the diagonal and for equal values), evaluated as e^{a_j} expm1(d) / d
    with d = a_i - a_j, so nearly equal eigenvalues lose no digits."""
    ai = a[..., :, None]
    aj = a[..., None, :]
    d = ai - aj
    with np.errstate(invalid="ignore", divide="ignore"):
        q = np.where(d == 0.0, 1.0, np.expm1(d) / np.where(d == 0.0, 1.0, d))
    return np.exp(aj) * q


def dlogl_dM(U, D, Uinv, mu, blv, rates, dP) -> np.ndarray:
    """d logL / dM (S x S) for M = mu Q, from dP [N-1, R, S, S] = d logL /
    d P_{v,r} and the eigensystem the likelihood was evaluated with."""
    U = np.asarray(U, dtype=np.float64)
    Ui = np.asarray(Uinv, dtype=np.float64)
    D = np.asarray(D, dtype=np.float64)
    tau = np.asarray(blv, dtype=np.float64)[:, None] * np.asarray(rates, dtype=np.float64)[None, :]
    a = tau[..., None] * (float(mu) * D)[None, None, :]
    Phi = tau[..., None, None] * _exp_divided_difference(a)
    X = U.T[None, None] @ np.asarray(dP, dtype=np.float64) @ Ui.T[None, None]
    return Ui.T @ np.sum(X * Phi, axis=(0, 1)) @ U.T


def gtr_gradient(base_freq, substitution_rates, U, D, Uinv, mu, blv, rates, dP, dpi):
    """(d logL / d substitution_rates [S(S-1)/2, setmatrix order],
    d logL / d base_freq [S]) of the GTR model (`SubstitutionModels.jl:62-74`):
    Q[i,j] = R[i,j] pi[j] (i != j), Q[i,i] = -sum_j Q[i,j], mu = 1 / sum of the
    row sums; pi also weights the root term.  The frequencies are taken as
    given (no renormalisation), as the reference's GTR takes them."""
    pi = np.asarray(base_freq, dtype=np.float64)
    S = pi.shape[0]
    Rm = setmatrix(substitution_rates)
    Q = Rm * pi[None, :]
    Q[np.diag_indices(S)] = -Q.sum(axis=1)
    G = dlogl_dM(U, D, Uinv, mu, blv, rates, dP)
    gQ = float(np.sum(G * Q))          # <G, Q>: the mu channel
    mu = float(mu)
    d_exch = []
    for j in range(S):
        for i in range(j):
            dq = G[i, j] * pi[j] + G[j, i] * pi[i] - G[i, i] * pi[j] - G[j, j] * pi[i]
            dmu = -mu * mu * (pi[i] + pi[j])
            d_exch.append(dmu * gQ + mu * dq)
    d_pi = np.zeros(S)
    for j in range(S):
        dq = sum(Rm[i, j] * (G[i, j] - G[i, i]) for i in range(S) if i != j)
        dmu = -mu * mu * sum(Rm[i, j] for i in range(S) if i != j)
        d_pi[j] = dmu * gQ + mu * dq + float(np.asarray(dpi)[j])
    return np.array(d_exch), d_pi


def restriction_gradient(base_freq, U, D, Uinv, mu, blv, rates, dP, dpi):
    """d logL / d base_freq [2] of the restriction model
    (`SubstitutionModels.jl:13-24`): its U (hence Q = [[-p2, p2], [1-p2,
    -(1-p2)]]) depends on pi[2] only, mu = 1 / (2 pi[1] pi[2]) on both, and pi
    weights the root term."""
    f = np.asarray(base_freq, dtype=np.float64)
    Q = np.array([[-f[1], f[1]], [1.0 - f[1], -(1.0 - f[1])]])
    G = dlogl_dM(U, D, Uinv, mu, blv, rates, dP)
    gQ = float(np.sum(G * Q))
    mu = float(mu)
    dQ2 = np.array([[-1.0, 1.0], [-1.0, 1.0]])
    d1 = (-mu / f[0]) * gQ + float(dpi[0])
    d2 = (-mu / f[1]) * gQ + mu * float(np.sum(G * dQ2)) + float(dpi[1])
    return np.array([d1, d2])


def rate_gradient(U, D, Uinv, mu, blv, rates, dP) -> np.ndarray:
    """d logL / d rho_r [R] of the category rates from dP [N-1, R, S, S]:
    sum_v <A_{v,r}, dP_{v,r}/d rho_r>, dP/d rho = U diag(mu t D e^{a}) Uinv."""
    U = np.asarray(U, dtype=np.float64)
    Ui = np.asarray(Uinv, dtype=np.float64)
    D = np.asarray(D, dtype=np.float64)
    t = np.asarray(blv, dtype=np.float64)
    rho = np.asarray(rates, dtype=np.float64)
    a = (t[:, None] * rho[None, :])[..., None] * (float(mu) * D)[None, None, :]
    scale = (float(mu) * t[:, None, None] * D[None, None, :]) * np.exp(a)
    dPdrho = (U[None, None] * scale[..., None, :]) @ Ui[None, None]
    return np.sum(np.asarray(dP, dtype=np.float64) * dPdrho, axis=(0, 2, 3))
