"""ctypes front end of the CPU oracle (liboracle.so, built from
felsenstein_oracle.c by the Makefile next to this file).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the CPU baseline and
never as the product path.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        lib = ctypes.CDLL(_LIB)
        lib.oracle_logl_grad.restype = ctypes.c_int
        lib.oracle_logl_grad.argtypes = [ctypes.c_int] * 4 + [_ip] * 4 + [_dp] * 5 + [
            ctypes.c_double] + [_dp] * 5
        lib.oracle_transition.restype = None
        lib.oracle_transition.argtypes = [ctypes.c_int] * 3 + [_dp] * 3 + [
            ctypes.c_double, _dp, _dp, ctypes.c_int, _dp]
        lib.oracle_set_threads.argtypes = [ctypes.c_int]
        lib.oracle_get_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip)


def set_threads(n: int) -> None:
    load().oracle_set_threads(int(n))


def get_threads() -> int:
    return int(load().oracle_get_threads())


def transition(U, D, Uinv, mu, rates, blv, deriv=False) -> np.ndarray:
    """P (or dP/dt) with shape (S, S, R, NB), Fortran order."""
    U = np.asfortranarray(U, dtype=np.float64)
    Uinv = np.asfortranarray(Uinv, dtype=np.float64)
    D = np.ascontiguousarray(D, dtype=np.float64)
    rates = np.ascontiguousarray(rates, dtype=np.float64)
    blv = np.ascontiguousarray(blv, dtype=np.float64)
    S, R, NB = U.shape[0], rates.shape[0], blv.shape[0]
    out = np.zeros((S, S, R, NB), order="F")
    load().oracle_transition(S, R, NB, _d(U), _d(D), _d(Uinv), float(mu), _d(rates),
                             _d(blv), 1 if deriv else 0, _d(out))
    return out


def logl_grad(ta, x, U, D, Uinv, mu, rates, pi, blv, weights=None, grad=True):
    """(ll, grad[N-1]) for a tree given as tree.TreeArrays `ta` and tip
    partials x of shape (S, C, N) (any memory order)."""
    x = np.asfortranarray(x, dtype=np.float64)
    S, C, N = x.shape
    if N != ta.N:
        raise ValueError("x has the wrong number of nodes")
    U = np.asfortranarray(U, dtype=np.float64)
    Uinv = np.asfortranarray(Uinv, dtype=np.float64)
    D = np.ascontiguousarray(D, dtype=np.float64)
    rates = np.ascontiguousarray(rates, dtype=np.float64)
    pi = np.ascontiguousarray(pi, dtype=np.float64)
    blv = np.ascontiguousarray(blv, dtype=np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    po = np.ascontiguousarray(ta.postorder, dtype=np.int32)
    par = np.ascontiguousarray(ta.parent, dtype=np.int32)
    cp = np.ascontiguousarray(ta.child_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(ta.child_idx, dtype=np.int32)
    ll = np.zeros(1)
    g = np.zeros(N - 1) if grad else None
    st = load().oracle_logl_grad(N, S, C, rates.shape[0], _i(po), _i(par), _i(cp), _i(ci),
                                 _d(x), _d(w), _d(U), _d(D), _d(Uinv), float(mu), _d(rates),
                                 _d(pi), _d(blv), _d(ll), _d(g))
    if st != 0:
        raise MemoryError("oracle could not allocate its arrays")
    return (float(ll[0]), g) if grad else float(ll[0])


def compound_dirichlet(blv, internal, alpha, a, beta, c):
    """(logprior, gradient) of CompoundDirichlet(alpha, a, beta, c) at branch
    lengths blv (internal_logpdf, src/Likelihood/Prior.jl:1-37; its gradient
    in closed form).  internal[i] == 1 marks the branch above an internal
    node."""
    import math
    blv = np.asarray(blv, dtype=np.float64)
    inl = np.asarray(internal) == 1
    T = 0.0
    s_leaf = s_int = 0.0
    nterm = 0.0
    for i in range(blv.shape[0]):
        T += blv[i]
        if inl[i]:
            s_int += math.log(blv[i])
        else:
            s_leaf += math.log(blv[i])
            nterm += 1.0
    n_int = nterm - 3.0
    first = alpha * math.log(beta) - math.lgamma(alpha) - T * beta
    second = -math.lgamma(a) - math.lgamma(c) + math.lgamma(a + c)
    third = s_leaf * (a - 1.0) + s_int * (a * c - 1.0)
    k = alpha - a * nterm - a * c * n_int
    fourth = k * math.log(T)
    g = np.array([-beta + ((a * c - 1.0) if inl[i] else (a - 1.0)) / blv[i] + k / T
                  for i in range(blv.shape[0])])
    return first + second + third + fourth, g
