"""CPU oracle of the learned surrogate (test infrastructure only: imported by
tests/, __graft_entry__.smoke() and bench.py's checks, never by the product
path).

The surrogate has no counterpart in MCPhylo.jl; this module restates the
network and its gradients in numpy so the device kernels can be checked:

    z  = (log x - x_mean) / x_sd                       (features)
    h1 = tanh(W1 z + b1);  h2 = tanh(W2 h1 + b2)
    y  = w3.h2 + b3[0] + v.z + q.(z*z);   logL~ = y y_sd + y_mean
    loss = (1/2n) sum_r (y_r - t_r)^2 + (lam / 2nd) sum_{r, c<d} (g_rc - G_rc)^2
    t = (logL - y_mean) / y_sd,  g = dy/dz (input gradient),
    G = dlogL/dx * x * x_sd / y_sd (standardised exact gradient)

Two modes:
  exact  -- every product in float64 (the mathematical network);
  bf16   -- the device's arithmetic: GEMM operands rounded to bfloat16
            (round to nearest even), products accumulated in float64, and the
            arrays the device stores in FP32 rounded to float32 at the same
            points, so only FP32 accumulation order separates it from the
            device.
Pinned by central finite differences of the exact mode (tests/test_surrogate
_oracle.py): parameter gradients of the loss (with and without the
gradient-matching term) and input gradients of logL~.
"""

from __future__ import annotations

import numpy as np

BLOCKS = ("W1", "b1", "W2", "b2", "w3", "b3", "v", "q")


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def unflatten(flat, hidden, D):
    """Blocks of the flat parameter layout W1 [H][D] | b1 [H] | W2 [H][H] |
    b2 [H] | w3 [H] | b3 [64] | v [D] | q [D]."""
    H = hidden
    sizes = [H * D, H, H * H, H, H, 64, D, D]
    shapes = [(H, D), (H,), (H, H), (H,), (H,), (64,), (D,), (D,)]
    out, o = {}, 0
    for name, n, sh in zip(BLOCKS, sizes, shapes):
        out[name] = np.asarray(flat[o:o + n]).reshape(sh)
        o += n
    if o != len(flat):
        raise ValueError("flat parameter vector has the wrong length")
    return out


def features(X, x_mean, x_sd, D, dtype=np.float32):
    """Standardised features padded to D columns: float32 as the device
    stores them, or float64 for the exact network."""
    X = np.atleast_2d(np.asarray(X, dtype=np.float64))
    n, d = X.shape
    Z = np.zeros((n, D), dtype=dtype)
    Z[:, :d] = ((np.log(X) - np.asarray(x_mean, np.float32).astype(np.float64))
                / np.asarray(x_sd, np.float32).astype(np.float64)).astype(dtype)
    return Z


class MLPOracle:
    def __init__(self, flat, hidden, D, mode="exact"):
        if mode not in ("exact", "bf16"):
            raise ValueError(mode)
        self.mode = mode
        self.p = {k: v.astype(np.float64) for k, v in unflatten(flat, hidden, D).items()}
        self.p32 = unflatten(np.asarray(flat, np.float32), hidden, D)
        self.H, self.D = hidden, D

    # matrix product with the mode's operand rounding
    def _mm(self, a, b):
        if self.mode == "bf16":
            return bf16(a).astype(np.float64) @ bf16(b).astype(np.float64)
        return np.asarray(a, np.float64) @ np.asarray(b, np.float64)

    def _store(self, a):
        return a.astype(np.float32) if self.mode == "bf16" else a

    def forward(self, Z):
        p = self.p
        P = self.p32 if self.mode == "bf16" else p
        Z64 = Z.astype(np.float64)
        h1 = self._store(np.tanh(self._mm(Z, P["W1"].T) + p["b1"]))
        h2 = self._store(np.tanh(self._mm(h1, P["W2"].T) + p["b2"]))
        h2d = np.asarray(h2, np.float64)
        y = h2d @ p["w3"] + p["b3"][0] + Z64 @ p["v"] + (Z64 * Z64) @ p["q"]
        return h1, h2, self._store(y)

    def loss_grad(self, Z, t):
        """(loss, flat gradient in float64) of (1/2n) sum (y - t)^2."""
        p = self.p
        P = self.p32 if self.mode == "bf16" else p
        n = Z.shape[0]
        h1, h2, y = self.forward(Z)
        h1d, h2d = np.asarray(h1, np.float64), np.asarray(h2, np.float64)
        Z64 = Z.astype(np.float64)
        e = np.asarray(y, np.float64) - t
        dy = self._store(e / n)
        dyd = np.asarray(dy, np.float64)
        dA2 = self._store(dyd[:, None] * p["w3"][None, :] * (1.0 - h2d * h2d))
        g = {
            "w3": dyd @ h2d,
            "v": dyd @ Z64,
            "q": dyd @ (Z64 * Z64),
            "b3": np.concatenate([[dyd.sum()], np.zeros(63)]),
            "b2": np.asarray(dA2, np.float64).sum(axis=0),
            "W2": self._mm(np.asarray(dA2).T, h1),
        }
        dH1 = self._store(self._mm(dA2, P["W2"]) * (1.0 - h1d * h1d))
        g["b1"] = np.asarray(dH1, np.float64).sum(axis=0)
        g["W1"] = self._mm(np.asarray(dH1).T, Z)
        loss = 0.5 * float(e @ e) / n
        return loss, np.concatenate([np.asarray(g[k], np.float64).reshape(-1) for k in BLOCKS])

    def loss_grad_sobolev(self, Z, t, GZ, lam, d):
        """(loss, flat gradient) with the gradient-matching term: a second
        backward pass through the input-gradient graph
        (delta2 = w3 s2, T1 = delta2 W2, delta1 = T1 s1, g = delta1 W1 + v + 2qz)."""
        p = self.p
        P = self.p32 if self.mode == "bf16" else p
        st, mm = self._store, self._mm
        n = Z.shape[0]
        h1, h2, y = self.forward(Z)
        h1d, h2d = np.asarray(h1, np.float64), np.asarray(h2, np.float64)
        s1, s2 = 1.0 - h1d * h1d, 1.0 - h2d * h2d
        Z64 = Z.astype(np.float64)
        e = np.asarray(y, np.float64) - t
        dyd = np.asarray(st(e / n), np.float64)
        D2 = st(p["w3"][None, :] * s2)
        T1 = st(mm(D2, P["W2"]))
        Dl1 = st(np.asarray(T1, np.float64) * s1)
        dZ = np.asarray(st(mm(Dl1, P["W1"])), np.float64)
        gin = dZ + p["v"][None, :] + 2.0 * p["q"][None, :] * Z64
        mask = np.zeros(Z.shape[1])
        mask[:d] = 1.0
        diff = (gin - np.asarray(GZ, np.float64)) * mask[None, :]
        lam_n = lam / (n * d)
        rho = st(lam_n * diff)
        rhod = np.asarray(rho, np.float64)
        K1 = np.asarray(st(mm(rho, P["W1"].T)), np.float64)
        mu1 = st(K1 * s1)
        gam1 = -2.0 * h1d * K1 * np.asarray(T1, np.float64)
        K2 = np.asarray(st(mm(mu1, P["W2"].T)), np.float64)
        D2d = np.asarray(D2, np.float64)
        A2 = st(dyd[:, None] * D2d - 2.0 * h2d * K2 * p["w3"][None, :] * s2)
        g = {
            "w3": (dyd[:, None] * h2d + K2 * s2).sum(axis=0),
            "v": (dyd[:, None] * Z64 + rhod).sum(axis=0),
            "q": (dyd[:, None] * Z64 * Z64 + 2.0 * Z64 * rhod).sum(axis=0),
            "b3": np.concatenate([[dyd.sum()], np.zeros(63)]),
            "b2": np.asarray(A2, np.float64).sum(axis=0),
            "W2": mm(np.asarray(A2).T, h1) + mm(np.asarray(D2).T, mu1),
        }
        A1 = st((mm(A2, P["W2"]) + gam1) * s1)
        g["b1"] = np.asarray(A1, np.float64).sum(axis=0)
        g["W1"] = mm(np.asarray(A1).T, Z) + mm(np.asarray(Dl1).T, rho)
        loss = 0.5 * float(e @ e) / n + 0.5 * lam_n * float(np.sum(diff * diff))
        return loss, np.concatenate([np.asarray(g[k], np.float64).reshape(-1) for k in BLOCKS])

    def input_grad(self, Z):
        """d y / d z per row (float64)."""
        p = self.p
        P = self.p32 if self.mode == "bf16" else p
        h1, h2, _ = self.forward(Z)
        h1d, h2d = np.asarray(h1, np.float64), np.asarray(h2, np.float64)
        dA2 = self._store(p["w3"][None, :] * (1.0 - h2d * h2d))
        dH1 = self._store(self._mm(dA2, P["W2"]) * (1.0 - h1d * h1d))
        dZ = self._store(self._mm(dH1, P["W1"]))
        Z64 = Z.astype(np.float64)
        return np.asarray(dZ, np.float64) + p["v"][None, :] + 2.0 * p["q"][None, :] * Z64


def predict(flat, hidden, D, X, x_mean, x_sd, y_mean, y_sd, mode="exact"):
    """(logL~ [n], d logL~/dx [n][d]) of the rows of X."""
    X = np.atleast_2d(np.asarray(X, dtype=np.float64))
    d = X.shape[1]
    o = MLPOracle(flat, hidden, D, mode)
    Z = features(X, x_mean, x_sd, D, np.float32 if mode == "bf16" else np.float64)
    _, _, y = o.forward(Z)
    gz = o.input_grad(Z)[:, :d]
    sd = np.asarray(x_sd, np.float32).astype(np.float64)
    ys = float(np.float32(y_sd))
    ll = np.asarray(y, np.float64) * ys + float(np.float32(y_mean))
    return ll, ys * gz / (sd[None, :] * X)


def param_grad(flat, hidden, D, X, logl, x_mean, x_sd, y_mean, y_sd, mode="exact",
               dlogl=None, grad_weight=0.0):
    """(loss, flat gradient) of the training loss over the rows of (X, logl)
    and, with grad_weight > 0, the exact gradients dlogl [n][d]."""
    X = np.atleast_2d(np.asarray(X, dtype=np.float64))
    Z = features(X, x_mean, x_sd, D, np.float32 if mode == "bf16" else np.float64)
    ym, ys = float(np.float32(y_mean)), float(np.float32(y_sd))
    t = (np.asarray(logl, np.float64) - ym) / ys
    if mode == "bf16":
        t = t.astype(np.float32).astype(np.float64)
    o = MLPOracle(flat, hidden, D, mode)
    if grad_weight <= 0.0:
        return o.loss_grad(Z, t)
    n, d = X.shape
    GZ = np.zeros((n, D))
    GZ[:, :d] = (np.asarray(dlogl, np.float64) * X
                 * np.asarray(x_sd, np.float32).astype(np.float64) / ys)
    if mode == "bf16":
        GZ = GZ.astype(np.float32).astype(np.float64)
    return o.loss_grad_sobolev(Z, t, GZ, float(np.float32(grad_weight)), d)
