"""CPU: the host chain rule from d logL / dP (and d logL / d pi at the root)
to the substitution model's parameters (models.dlogl_dM, gtr_gradient,
restriction_gradient), pinned by central finite differences on a likelihood
whose d logL / dP is known in closed form: a two-leaf tree (a cherry), where
logL = sum_{c,r} w_c log sum_s pi_s P1_r[s][x1_c] P2_r[s][x2_c]
(the root term of VectorizedFunctions.jl:76-87 with two leaf children)."""

import numpy as np
import pytest
from scipy.linalg import expm

from conftest import mcp

M = mcp.models


def _transition(U, D, Uinv, mu, t, rho):
    return (U * np.exp(mu * t * D * rho)[None, :]) @ Uinv


def _cherry(model, pi, exch, blv, rates, x1, x2, w):
    """(logL, dP [2, R, S, S], dpi [S]) of the cherry, A in closed form."""
    U, D, Uinv, mu = model(pi, exch)
    R, S = len(rates), len(pi)
    ll = 0.0
    dP = np.zeros((2, R, S, S))
    dpi = np.zeros(S)
    for r, rho in enumerate(rates):
        P1 = _transition(U, D, Uinv, mu, blv[0], rho)
        P2 = _transition(U, D, Uinv, mu, blv[1], rho)
        for c in range(len(w)):
            a, b = P1[:, x1[c]], P2[:, x2[c]]
            L = float(np.sum(pi * a * b))
            ll += w[c] * np.log(L)
            dP[0, r, :, x1[c]] += w[c] * pi * b / L
            dP[1, r, :, x2[c]] += w[c] * pi * a / L
            dpi += w[c] * a * b / L
    return ll, dP, dpi


def _fd(f, x, h_rel=1e-6):
    x = np.asarray(x, dtype=np.float64)
    g = np.zeros_like(x)
    for k in range(x.size):
        h = h_rel * max(1.0, abs(x[k]))
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        g[k] = (f(xp) - f(xm)) / (2 * h)
    return g


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gtr_gradient_matches_fd_on_cherry(seed):
    rng = np.random.default_rng(seed)
    S, R, C = 4, 3, 40
    pi = rng.dirichlet(np.full(S, 5.0))
    exch = rng.exponential(1.0, S * (S - 1) // 2)
    blv = rng.exponential(0.3, 2)
    rates = np.array([0.3, 1.0, 1.7])[:R]
    x1, x2 = rng.integers(0, S, C), rng.integers(0, S, C)
    w = rng.integers(1, 4, C).astype(float)
    ll, dP, dpi = _cherry(M.GTR, pi, exch, blv, rates, x1, x2, w)
    U, D, Uinv, mu = M.GTR(pi, exch)
    d_exch, d_pi = M.gtr_gradient(pi, exch, U, D, Uinv, mu, blv, rates, dP, dpi)
    fd_exch = _fd(lambda e: _cherry(M.GTR, pi, e, blv, rates, x1, x2, w)[0], exch)
    fd_pi = _fd(lambda p: _cherry(M.GTR, p, exch, blv, rates, x1, x2, w)[0], pi)
    scale = max(1.0, np.max(np.abs(fd_exch)))
    assert np.max(np.abs(d_exch - fd_exch)) <= 1e-6 * scale, (d_exch, fd_exch)
    scale = max(1.0, np.max(np.abs(fd_pi)))
    assert np.max(np.abs(d_pi - fd_pi)) <= 1e-6 * scale, (d_pi, fd_pi)
    # the branch-length gradient is the same contraction with dP/dt
    for v in range(2):
        gv = sum(np.sum(dP[v, r] * ((U * (mu * D * rho * np.exp(mu * blv[v] * D * rho))[None, :])
                                    @ Uinv)) for r, rho in enumerate(rates))
        bp, bm = blv.copy(), blv.copy()
        bp[v] += 1e-6
        bm[v] -= 1e-6
        fd = (_cherry(M.GTR, pi, exch, bp, rates, x1, x2, w)[0]
              - _cherry(M.GTR, pi, exch, bm, rates, x1, x2, w)[0]) / 2e-6
        assert abs(gv - fd) <= 1e-6 * max(1.0, abs(fd))


def test_restriction_gradient_matches_fd_on_cherry():
    rng = np.random.default_rng(7)
    C, rates = 30, np.array([0.5, 1.5])
    pi = np.array([0.35, 0.65])
    blv = np.array([0.2, 0.45])
    x1, x2 = rng.integers(0, 2, C), rng.integers(0, 2, C)
    w = np.ones(C)
    ll, dP, dpi = _cherry(M.Restriction, pi, None, blv, rates, x1, x2, w)
    U, D, Uinv, mu = M.Restriction(pi)
    d_pi = M.restriction_gradient(pi, U, D, Uinv, mu, blv, rates, dP, dpi)
    fd = _fd(lambda p: _cherry(M.Restriction, p, None, blv, rates, x1, x2, w)[0], pi)
    assert np.max(np.abs(d_pi - fd)) <= 1e-6 * max(1.0, np.max(np.abs(fd))), (d_pi, fd)


@pytest.mark.parametrize("equal", [False, True])
def test_dlogl_dM_matches_matrix_exponential_fd(equal):
    """<A, d exp(tau M)> against finite differences of scipy's expm, also
    with repeated eigenvalues (a JC-like Q), where the divided difference of
    exp takes its limit."""
    rng = np.random.default_rng(11)
    S = 4
    pi = np.full(S, 0.25) if equal else rng.dirichlet(np.full(S, 4.0))
    exch = np.ones(6) if equal else rng.exponential(1.0, 6)
    U, D, Uinv, mu = M.GTR(pi, exch)
    Q = (U * D[None, :]) @ Uinv
    blv, rates = np.array([0.3, 0.8]), np.array([0.5, 2.0])
    A = rng.normal(size=(2, 2, S, S))
    G = M.dlogl_dM(U, D, Uinv, mu, blv, rates, A)
    E = rng.normal(size=(S, S))

    def f(h):
        Mh = mu * Q + h * E
        return sum(np.sum(A[v, r] * expm(blv[v] * rates[r] * Mh))
                   for v in range(2) for r in range(2))

    fd = (f(1e-6) - f(-1e-6)) / 2e-6
    assert abs(np.sum(G * E) - fd) <= 1e-7 * max(1.0, abs(fd)), (np.sum(G * E), fd)


def test_exp_divided_difference_limits():
    a = np.array([-1.0, -1.0, 0.0, -1.0 + 1e-12])
    F = M._exp_divided_difference(a)
    assert np.allclose(np.diag(F), np.exp(a), rtol=1e-15, atol=0)
    assert F[0, 1] == np.exp(-1.0)
    assert abs(F[0, 3] - np.exp(-1.0)) <= 1e-11
    assert abs(F[0, 2] - (np.exp(-1.0) - 1.0) / (-1.0)) <= 1e-15
