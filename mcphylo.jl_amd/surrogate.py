"""Learned surrogate of the pruning log-likelihood (delayed-acceptance first
stage), over the mcp_mlp_* entry points of libmcphylo_hip.so.

The network maps standardised log branch lengths z (one input per branch,
i.e. per split of the fixed topology) to the standardised log-likelihood:

    h1 = tanh(W1 z + b1); h2 = tanh(W2 h1 + b2)
    y  = w3.h2 + b3 + v.z + q.(z*z);   logL~ = y * y_sd + y_mean

Training (Adam on the mean-squared error), inference and the surrogate
leapfrog run on the GPU: BF16 MFMA matrix products with FP32 accumulation.
The reference has no such component: MCPhylo evaluates the exact likelihood
at every leapfrog (src/samplers/pnuts.jl:300-330); this module serves the
delayed-acceptance sampler in `sampler.run_da_nuts`, which keeps the exact
FP64 likelihood as the second stage.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, dptr, fptr

__all__ = ["Surrogate", "BLOCKS"]

# parameter blocks in the flat layout
BLOCKS = ("W1", "b1", "W2", "b2", "w3", "b3", "v", "q")


def _rows(X, d):
    X = np.ascontiguousarray(X, dtype=np.float64)
    if X.ndim == 1:
        X = X.reshape(1, -1)
    if X.ndim != 2 or X.shape[1] != d:
        raise ValueError(f"expected rows of {d} branch lengths, got shape {X.shape}")
    return X


class Surrogate:
    """Device MLP d -> hidden -> hidden -> 1 with a separable quadratic skip."""

    def __init__(self, n_inputs: int, hidden: int = 256, device: int = 0, seed: int = 0):
        lib = _lib.load_library()
        self._h = ctypes.c_void_p()
        check(lib.mcp_mlp_create(ctypes.byref(self._h), int(n_inputs), int(hidden), int(device),
                                 int(seed)))
        self.d = int(n_inputs)
        self.hidden = int(hidden)
        off = np.zeros(9, dtype=np.int64)
        check(lib.mcp_mlp_layout(self._h, off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        self.offsets = off
        self.n_params = int(off[8])
        self.D = int((off[1] - off[0]) // self.hidden)
        self.x_mean = np.zeros(self.d, dtype=np.float32)
        self.x_sd = np.ones(self.d, dtype=np.float32)
        self.y_mean, self.y_sd = 0.0, 1.0

    # -- lifetime ------------------------------------------------------------
    def close(self):
        if self._h and self._h.value:
            _lib.load_library().mcp_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- parameters ------------------------------------------------------------
    def shapes(self):
        H, D = self.hidden, self.D
        return {"W1": (H, D), "b1": (H,), "W2": (H, H), "b2": (H,), "w3": (H,), "b3": (64,),
                "v": (D,), "q": (D,)}

    def get_flat(self) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        check(_lib.load_library().mcp_mlp_get_params(self._h, fptr(out)))
        return out

    def set_flat(self, flat) -> None:
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        if flat.shape != (self.n_params,):
            raise ValueError(f"expected {self.n_params} parameters")
        check(_lib.load_library().mcp_mlp_set_params(self._h, fptr(flat)))

    def unflatten(self, flat) -> dict:
        """Views of the blocks of a flat parameter (or gradient) vector."""
        sh = self.shapes()
        return {name: flat[self.offsets[i]:self.offsets[i + 1]].reshape(sh[name])
                for i, name in enumerate(BLOCKS)}

    def params(self) -> dict:
        return self.unflatten(self.get_flat())

    # -- scaling and data ----------------------------------------------------
    def set_scaling(self, x_mean, x_sd, y_mean: float, y_sd: float) -> None:
        self.x_mean = np.ascontiguousarray(x_mean, dtype=np.float32).reshape(self.d)
        self.x_sd = np.ascontiguousarray(x_sd, dtype=np.float32).reshape(self.d)
        self.y_mean, self.y_sd = float(np.float32(y_mean)), float(np.float32(y_sd))
        check(_lib.load_library().mcp_mlp_set_scaling(self._h, fptr(self.x_mean), fptr(self.x_sd),
                                                      self.y_mean, self.y_sd))

    def fit_scaling(self, X, logl, sd_floor: float = 1e-3) -> None:
        """Standardise with the mean and sd of log X (per branch) and of logl."""
        L = np.log(_rows(X, self.d))
        y = np.asarray(logl, dtype=np.float64)
        self.set_scaling(L.mean(axis=0), np.maximum(L.std(axis=0), sd_floor), y.mean(),
                         max(float(y.std()), sd_floor))

    def set_data(self, X, logl, grad=None) -> None:
        """Resident training set: branch lengths X [n][d], logL [n] and
        optionally the exact gradients d logL / dx [n][d]."""
        X = _rows(X, self.d)
        y = np.ascontiguousarray(logl, dtype=np.float64).reshape(-1)
        if y.shape[0] != X.shape[0]:
            raise ValueError("X and logl have different numbers of rows")
        G = None if grad is None else _rows(grad, self.d)
        if G is not None and G.shape != X.shape:
            raise ValueError("grad must have the shape of X")
        check(_lib.load_library().mcp_mlp_set_data(self._h, dptr(X), dptr(y), dptr(G),
                                                   X.shape[0]))

    def train(self, epochs: int, batch: int = 512, lr: float = 1e-3, lr_final: float = 1e-5,
              seed: int = 0, grad_weight: float = 0.0) -> np.ndarray:
        """Adam epochs over the resident data (grad_weight > 0 adds gradient
        matching); returns the mean batch loss per epoch (standardised units)."""
        losses = np.zeros(max(1, int(epochs)), dtype=np.float32)
        check(_lib.load_library().mcp_mlp_train(self._h, int(batch), int(epochs), float(lr),
                                                float(lr_final), float(grad_weight), int(seed),
                                                fptr(losses)))
        return losses[:int(epochs)].astype(np.float64)

    def param_grad(self, X, logl, grad=None, grad_weight: float = 0.0):
        """(flat gradient of the training loss, loss) over the rows of X."""
        X = _rows(X, self.d)
        y = np.ascontiguousarray(logl, dtype=np.float64).reshape(-1)
        G = None if grad is None else _rows(grad, self.d)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        check(_lib.load_library().mcp_mlp_param_grad(self._h, dptr(X), dptr(y), dptr(G),
                                                     X.shape[0], float(grad_weight), fptr(g),
                                                     fptr(loss)))
        return g, float(loss[0])

    # -- inference ---------------------------------------------------------------
    def predict(self, X, grad: bool = False):
        """logL~ of each row (and d logL~/dx when grad)."""
        X = _rows(X, self.d)
        n = X.shape[0]
        ll = np.zeros(n)
        g = np.zeros((n, self.d)) if grad else None
        check(_lib.load_library().mcp_mlp_predict(self._h, dptr(X), n, dptr(ll), dptr(g)))
        return (ll, g) if grad else ll

    def use_graph(self, on: bool = True) -> None:
        check(_lib.load_library().mcp_mlp_use_graph(self._h, 1 if on else 0))

    def train_graph(self, on: bool = True) -> None:
        """Replay one captured graph per full training minibatch (default)."""
        check(_lib.load_library().mcp_mlp_train_graph(self._h, 1 if on else 0))

    def leapfrog(self, eps_delta, x, r, g, prior_scale: float | None = None) -> np.ndarray:
        """One surrogate leapfrog for every chain; x, r, g ([chains][d],
        C-contiguous float64) are updated in place; returns the log posterior."""
        for a in (x, r, g):
            if a.dtype != np.float64 or not a.flags.c_contiguous or a.size % self.d:
                raise ValueError("x, r, g must be C-contiguous float64 [chains][d]")
        chains = x.size // self.d
        ed = np.ascontiguousarray(eps_delta, dtype=np.float64)
        if ed.size != 2 * chains or r.size != x.size or g.size != x.size:
            raise ValueError("inconsistent leapfrog arrays")
        lp = np.zeros(chains)
        ps = -1.0 if prior_scale is None else float(prior_scale)
        check(_lib.load_library().mcp_mlp_leapfrog(self._h, chains, ps, dptr(ed), dptr(x), dptr(r),
                                                   dptr(g), dptr(lp)))
        return lp
