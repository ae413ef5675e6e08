"""Discrete-gamma category rates (Yang 1994), host side.

Restates `src/Likelihood/Rates.jl:11-58`, including its choices: the
chi-square has (2 alpha)/(2 beta) degrees of freedom, mean boundaries are
divided by 2 beta, median boundaries are not, and the mean method uses the
regularised lower incomplete gamma P(alpha+1, b*beta).
"""

from __future__ import annotations

import numpy as np
from scipy.special import gammainc
from scipy.stats import chi2

__all__ = ["discrete_gamma_rates", "mean_boundaries", "median_boundaries"]


def mean_boundaries(alpha: float, beta: float, k: int) -> np.ndarray:
    """`Rates.jl:42-49`; entry k is left unset in the reference (0 here)."""
    df = (2.0 * alpha) / (2.0 * beta)
    out = np.zeros(k)
    for i in range(1, k):
        out[i - 1] = chi2.ppf(i / k, df) / (2.0 * beta)
    return out


def median_boundaries(alpha: float, beta: float, k: int) -> np.ndarray:
    """`Rates.jl:51-58`."""
    df = (2.0 * alpha) / (2.0 * beta)
    return np.array([chi2.ppf(((i - 1) * 2 + 1) / (2.0 * k), df) for i in range(1, k + 1)])


def discrete_gamma_rates(alpha: float, beta: float, k: int, method: str = "mean") -> np.ndarray:
    """`Rates.jl:11-37`."""
    factor = alpha / beta * k
    if method == "median":
        m = median_boundaries(alpha, beta, k)
        return m * (factor / m.sum())
    m = mean_boundaries(alpha, beta, k)
    for i in range(k - 1):
        m[i] = gammainc(alpha + 1.0, m[i] * beta)
    m[k - 1] = 1.0
    for i in range(k - 1, 0, -1):
        m[i] -= m[i - 1]
        m[i] *= factor
    m[0] *= factor
    return m
