"""Python handles over the C ABI: `Plan` (one tree) and `Batch` (many
(tree, chain) items of uneven size).  Device memory, level schedules and the
evaluation sequence live in the native library; these classes only marshal
host arrays."""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, dptr, iptr

__all__ = ["Plan", "Batch", "Model"]


class Model:
    """Eigensystem of Q plus category rates and base frequencies."""

    __slots__ = ("U", "D", "Uinv", "mu", "rates", "pi")

    def __init__(self, U, D, Uinv, mu, rates, pi):
        self.U = np.asfortranarray(U, dtype=np.float64)
        self.D = np.ascontiguousarray(D, dtype=np.float64)
        self.Uinv = np.asfortranarray(Uinv, dtype=np.float64)
        self.mu = float(mu)
        self.rates = np.ascontiguousarray(np.atleast_1d(rates), dtype=np.float64)
        self.pi = np.ascontiguousarray(pi, dtype=np.float64)
        S = self.D.shape[0]
        if self.U.shape != (S, S) or self.Uinv.shape != (S, S) or self.pi.shape != (S,):
            raise ValueError("model arrays have inconsistent shapes")

    def key(self) -> bytes:
        return (self.U.tobytes("F") + self.D.tobytes() + self.Uinv.tobytes("F") +
                np.float64(self.mu).tobytes() + self.rates.tobytes() + self.pi.tobytes())


class _Handle:
    _destroy = None

    def __init__(self):
        self._h = ctypes.c_void_p()

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h and self._h.value:
            getattr(_lib.load_library(), self._destroy)(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def use_graph(self, on: bool = True) -> None:
        check(_lib.load_library().mcp_use_graph(self._h, 1 if on else 0))

    def set_schedule(self, mode: str, hybrid_width: int | None = None) -> None:
        """'fused' (one launch per evaluation), 'levels' (one launch per tree
        level) or 'hybrid' (per-level launches for wide levels, one fused
        launch for the narrow top levels of at most hybrid_width nodes)."""
        m = {"fused": _lib.MCP_SCHEDULE_FUSED, "levels": _lib.MCP_SCHEDULE_LEVELS,
             "hybrid": _lib.MCP_SCHEDULE_HYBRID}[mode]
        if hybrid_width is not None:
            check(_lib.load_library().mcp_set_hybrid_width(self._h, int(hybrid_width)))
        check(_lib.load_library().mcp_set_schedule(self._h, m))

    def set_fused_waves(self, waves: int) -> None:
        """Wavefronts per block of the fused launches (1..8)."""
        check(_lib.load_library().mcp_set_fused_waves(self._h, int(waves)))

    def set_tiles(self, S: int, cpt: int) -> None:
        """Columns per thread of the level kernels for state count S (-1:
        built default; for S = 20, 0 = FP64-MFMA kernels)."""
        check(_lib.load_library().mcp_set_tiles(self._h, int(S), int(cpt)))

    def tiles(self, S: int) -> int:
        out = np.zeros(1, dtype=np.int32)
        check(_lib.load_library().mcp_get_tiles(self._h, int(S), iptr(out)))
        return int(out[0])

    def schedule(self):
        """(mode name, hybrid width) of the current launch schedule."""
        m = np.zeros(1, dtype=np.int32)
        wv = np.zeros(1, dtype=np.int32)
        check(_lib.load_library().mcp_get_schedule(self._h, iptr(m), iptr(wv)))
        return {0: "fused", 1: "levels", 2: "hybrid"}[int(m[0])], int(wv[0])

    def autotune_schedule(self, reps: int = 5):
        """Time levels / hybrid (widths 1..16) / fused and keep the fastest.
        Returns ((mode, width), [(mode, width, ms), ...])."""
        cap = 16
        rep = np.zeros(3 * cap)
        n = np.zeros(1, dtype=np.int32)
        check(_lib.load_library().mcp_autotune_schedule(self._h, int(reps), dptr(rep), cap,
                                                        iptr(n)))
        names = {0: "fused", 1: "levels", 2: "hybrid"}
        rows = [(names[int(rep[3 * k])], int(rep[3 * k + 1]), float(rep[3 * k + 2]))
                for k in range(min(int(n[0]), cap))]
        return self.schedule(), rows

    def launch_bytes(self) -> np.ndarray:
        """Algorithmic bytes of every launch of the last timed evaluation."""
        lib = _lib.load_library()
        n = lib.mcp_get_launch_bytes(self._h, None, 0)
        if n < 0:
            check(-n)
        out = np.zeros(n)
        lib.mcp_get_launch_bytes(self._h, dptr(out), n)
        return out

    def autotune(self, reps: int = 5):
        """Time every built tile width per state count on this engine's data
        and keep the fastest.  Returns ({S: cpt}, [(S, cpt, ms), ...])."""
        cap = 64
        rep = np.zeros(3 * cap)
        n = np.zeros(1, dtype=np.int32)
        check(_lib.load_library().mcp_autotune(self._h, int(reps), dptr(rep), cap, iptr(n)))
        rows = [(int(rep[3 * k]), int(rep[3 * k + 1]), float(rep[3 * k + 2]))
                for k in range(min(int(n[0]), cap))]
        chosen = {S: self.tiles(S) for S in sorted({r[0] for r in rows})}
        return chosen, rows

    def set_timing(self, on: bool = True) -> None:
        check(_lib.load_library().mcp_set_timing(self._h, 1 if on else 0))

    def timing(self) -> np.ndarray:
        """ms of the last evaluation: [total, P build, post-order, pre-order, reduce]."""
        out = np.zeros(5)
        check(_lib.load_library().mcp_get_timing(self._h, dptr(out)))
        return out

    def launch_times(self):
        """(ms, phase) of every kernel launch of the last timed evaluation."""
        lib = _lib.load_library()
        n = lib.mcp_get_launch_times(self._h, None, None, 0)
        if n < 0:
            check(-n)
        ms = np.zeros(max(n, 1))
        ph = np.zeros(max(n, 1), dtype=np.int32)
        lib.mcp_get_launch_times(self._h, dptr(ms), iptr(ph), n)
        return ms[:n], ph[:n]

    def set_prior_exponential(self, item: int, scale: float) -> None:
        """Exponential(mean=scale) prior on every branch of `item` for the leapfrog."""
        check(_lib.load_library().mcp_hmc_set_prior(self._h, int(item), 1, float(scale)))

    def set_prior_compound_dirichlet(self, item: int, alpha: float, a: float, beta: float,
                                     c: float) -> None:
        """CompoundDirichlet(alpha, a, beta, c) tree-length prior of `item` for
        the leapfrog (src/Likelihood/Prior.jl:1-37)."""
        p = np.array([alpha, a, beta, c], dtype=np.float64)
        check(_lib.load_library().mcp_hmc_set_prior_params(self._h, int(item), 2, dptr(p), 4))

    def leapfrog(self, eps_delta, x, r, g):
        """One device leapfrog for every item (see mcp_hmc_leapfrog).  x, r, g
        are updated in place (float64, blv_concat layout); returns
        (logpost[n_items], ll[n_items])."""
        ed = np.ascontiguousarray(eps_delta, dtype=np.float64)
        n = ed.shape[0] // 2
        lp = np.zeros(n)
        ll = np.zeros(n)
        for a in (x, r, g):
            if a.dtype != np.float64 or not a.flags.c_contiguous:
                raise ValueError("x, r, g must be contiguous float64 arrays")
        check(_lib.load_library().mcp_hmc_leapfrog(self._h, dptr(ed), dptr(x), dptr(r), dptr(g),
                                                   dptr(lp), dptr(ll)))
        return lp, ll

    def algorithmic_bytes(self) -> np.ndarray:
        out = np.zeros(2)
        check(_lib.load_library().mcp_get_bytes(self._h, dptr(out)))
        return out

    def sizes(self) -> tuple[int, int]:
        out = np.zeros(2, dtype=np.int64)
        check(_lib.load_library().mcp_get_sizes(
            self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return int(out[0]), int(out[1])


class Plan(_Handle):
    """Device plan for one tree (the PhyloDist drop-in).

    parent:  int32[N] parent num of node num i+1 (0 for the root, num N)
    x:       tip partials (S, C, N), reference layout (Fortran order)
    weights: optional pattern weights [C]
    """

    _destroy = "mcp_plan_destroy"

    def __init__(self, parent, x, n_rates: int, weights=None, device: int = 0,
                 compress: bool = False):
        super().__init__()
        self.parent = np.ascontiguousarray(parent, dtype=np.int32)
        x = np.asfortranarray(x, dtype=np.float64)
        S, C, N = x.shape
        if N != self.parent.shape[0]:
            raise _lib.McpDimensionMismatch(_lib.MCP_ERR_DIMENSION,
                                            "x and the tree disagree on the number of nodes")
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        if w is not None and w.shape != (C,):
            raise _lib.McpDimensionMismatch(_lib.MCP_ERR_DIMENSION, "weights must have length C")
        self.N, self.S, self.C, self.R = N, S, C, int(n_rates)
        d = _lib.PlanDesc(N, S, C, self.R, iptr(self.parent), dptr(x), dptr(w), int(device),
                          _lib.MCP_FLAG_COMPRESS_PATTERNS if compress else 0)
        lib = _lib.load_library()
        check(lib.mcp_plan_create(ctypes.byref(self._h), ctypes.byref(d)))
        self._model_key = None

    def set_model(self, m: Model) -> None:
        k = m.key()
        if k == self._model_key:
            return
        if m.D.shape[0] != self.S or m.rates.shape[0] != self.R:
            raise _lib.McpDimensionMismatch(_lib.MCP_ERR_DIMENSION,
                                            "model does not match the plan's S or R")
        check(_lib.load_library().mcp_set_model(self._h, dptr(m.U), dptr(m.D), dptr(m.Uinv), m.mu,
                                                dptr(m.rates), dptr(m.pi)))
        self._model_key = k

    def set_topology(self, parent) -> None:
        p = np.ascontiguousarray(parent, dtype=np.int32)
        check(_lib.load_library().mcp_set_topology(self._h, iptr(p)))
        self.parent = p

    def logl(self, blv) -> float:
        b = np.ascontiguousarray(blv, dtype=np.float64)
        if b.shape != (self.N - 1,):
            raise _lib.McpDimensionMismatch(_lib.MCP_ERR_DIMENSION, "blv must have length N-1")
        out = np.zeros(1)
        check(_lib.load_library().mcp_logl(self._h, dptr(b), dptr(out)))
        return float(out[0])

    def logl_grad(self, blv):
        b = np.ascontiguousarray(blv, dtype=np.float64)
        if b.shape != (self.N - 1,):
            raise _lib.McpDimensionMismatch(_lib.MCP_ERR_DIMENSION, "blv must have length N-1")
        out = np.zeros(1)
        g = np.zeros(self.N - 1)
        check(_lib.load_library().mcp_logl_grad(self._h, dptr(b), dptr(out), dptr(g)))
        return float(out[0]), g

    def logl_grad_by_rate(self, blv):
        """(ll, G [N-1, R]): per-(branch, category) gradient components;
        G.sum(axis=1) is the branch gradient."""
        b = np.ascontiguousarray(blv, dtype=np.float64)
        if b.shape != (self.N - 1,):
            raise _lib.McpDimensionMismatch(_lib.MCP_ERR_DIMENSION, "blv must have length N-1")
        out = np.zeros(1)
        g = np.zeros((self.N - 1, self.R))
        check(_lib.load_library().mcp_logl_grad_by_rate(self._h, dptr(b), dptr(out), dptr(g)))
        return float(out[0]), g

    def logl_grad_rates(self, blv, rates):
        """(ll, d ll / d rates [R]): gradient with respect to the category
        rates, sum_v t_v G[v, r] / rho_r."""
        ll, G = self.logl_grad_by_rate(blv)
        b = np.asarray(blv, dtype=np.float64)
        return ll, (b @ G) / np.asarray(rates, dtype=np.float64)

    def logl_grad_model(self, blv):
        """(ll, grad [N-1], dP [N-1, R, S, S], dpi [S]): besides the branch
        gradient, d ll / d P of every branch and category (row = parent
        state) and d ll / d pi through the root term (S <= 8)."""
        b = np.ascontiguousarray(blv, dtype=np.float64)
        if b.shape != (self.N - 1,):
            raise _lib.McpDimensionMismatch(_lib.MCP_ERR_DIMENSION, "blv must have length N-1")
        out = np.zeros(1)
        g = np.zeros(self.N - 1)
        dP = np.zeros((self.N - 1, self.R, self.S, self.S))
        dpi = np.zeros(self.S)
        check(_lib.load_library().mcp_logl_grad_model(self._h, dptr(b), dptr(out), dptr(g),
                                                      dptr(dP), dptr(dpi)))
        return float(out[0]), g, dP, dpi


class Batch(_Handle):
    """Ragged batch of items.

    datasets: list of dicts {leaf_nums [L], weights or None, compress bool} plus either
              x_leaves (S, C, L) tip partials, or tip_codes uint8 (L, C) with
              code_table (K, S)
    items:    list of dicts {dataset int, parent int32[N], n_rates int}
    """

    _destroy = "mcp_batch_destroy"

    def __init__(self, datasets, items, device: int = 0, waves_per_block: int = 0):
        super().__init__()
        self._keep = []
        dsd = (_lib.DatasetDesc * len(datasets))()
        self.dataset_S = []
        for k, d in enumerate(datasets):
            leaf = np.ascontiguousarray(d["leaf_nums"], dtype=np.int32)
            w = d.get("weights")
            w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
            flags = _lib.MCP_FLAG_COMPRESS_PATTERNS if d.get("compress") else 0
            if d.get("x_leaves") is not None:
                x = np.asfortranarray(d["x_leaves"], dtype=np.float64)
                S, C, L = x.shape
                codes = table = None
                K = 0
            else:
                codes = np.ascontiguousarray(d["tip_codes"], dtype=np.uint8)
                table = np.ascontiguousarray(d["code_table"], dtype=np.float64)
                L, C = codes.shape
                K, S = table.shape
                x = None
            if leaf.shape != (L,):
                raise _lib.McpDimensionMismatch(_lib.MCP_ERR_DIMENSION, "leaf_nums length")
            self._keep += [x, leaf, w, codes, table]
            cp = None if codes is None else codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
            dsd[k] = _lib.DatasetDesc(S, C, L, iptr(leaf), dptr(x), dptr(w), flags, cp, K,
                                      dptr(table))
            self.dataset_S.append(S)
        itd = (_lib.ItemDesc * len(items))()
        self.item_nodes = []
        self.item_S = []
        self.item_R = []
        for i, it in enumerate(items):
            p = np.ascontiguousarray(it["parent"], dtype=np.int32)
            self._keep.append(p)
            itd[i] = _lib.ItemDesc(int(it["dataset"]), p.shape[0], int(it["n_rates"]), iptr(p))
            self.item_nodes.append(p.shape[0])
            self.item_S.append(self.dataset_S[int(it["dataset"])])
            self.item_R.append(int(it["n_rates"]))
        bd = _lib.BatchDesc(len(datasets), dsd, len(items), itd, int(device), int(waves_per_block))
        check(_lib.load_library().mcp_batch_create(ctypes.byref(self._h), ctypes.byref(bd)))
        self._keep = []
        self.n_items = len(items)
        self.blv_off = np.concatenate([[0], np.cumsum([n - 1 for n in self.item_nodes])])
        self._model_keys = [None] * self.n_items

    def set_model(self, item: int, m: Model) -> None:
        k = m.key()
        if k == self._model_keys[item]:
            return
        if m.D.shape[0] != self.item_S[item] or m.rates.shape[0] != self.item_R[item]:
            raise _lib.McpDimensionMismatch(_lib.MCP_ERR_DIMENSION,
                                            "model does not match the item's S or R")
        check(_lib.load_library().mcp_batch_set_model(self._h, int(item), dptr(m.U), dptr(m.D),
                                                      dptr(m.Uinv), m.mu, dptr(m.rates),
                                                      dptr(m.pi)))
        self._model_keys[item] = k

    def set_topology(self, item: int, parent) -> None:
        p = np.ascontiguousarray(parent, dtype=np.int32)
        check(_lib.load_library().mcp_batch_set_topology(self._h, int(item), iptr(p)))

    def _blv(self, blv_concat):
        b = np.ascontiguousarray(blv_concat, dtype=np.float64)
        if b.shape != (int(self.blv_off[-1]),):
            raise _lib.McpDimensionMismatch(_lib.MCP_ERR_DIMENSION,
                                            "blv_concat has the wrong length")
        return b

    def logl(self, blv_concat) -> np.ndarray:
        b = self._blv(blv_concat)
        ll = np.zeros(self.n_items)
        check(_lib.load_library().mcp_batch_logl(self._h, dptr(b), dptr(ll)))
        return ll

    def logl_grad(self, blv_concat):
        b = self._blv(blv_concat)
        ll = np.zeros(self.n_items)
        g = np.zeros(b.shape[0])
        check(_lib.load_library().mcp_batch_logl_grad(self._h, dptr(b), dptr(ll), dptr(g)))
        return ll, g

    def logl_grad_by_rate(self, blv_concat):
        """(ll [items], [G_i [N_i-1, R_i]]) per-(branch, category) gradient
        components of every item."""
        b = self._blv(blv_concat)
        ll = np.zeros(self.n_items)
        sizes = [(n - 1) * r for n, r in zip(self.item_nodes, self.item_R)]
        g = np.zeros(int(np.sum(sizes)))
        check(_lib.load_library().mcp_batch_logl_grad_by_rate(self._h, dptr(b), dptr(ll),
                                                              dptr(g)))
        off = np.concatenate([[0], np.cumsum(sizes)])
        return ll, [g[off[i]:off[i + 1]].reshape(self.item_nodes[i] - 1, self.item_R[i])
                    for i in range(self.n_items)]

    def logl_grad_model(self, blv_concat):
        """(ll [items], grad_concat, [dP_i [N_i-1, R_i, S_i, S_i]], [dpi_i [S_i]]):
        the model gradient of every item (see Plan.logl_grad_model)."""
        b = self._blv(blv_concat)
        ll = np.zeros(self.n_items)
        g = np.zeros(b.shape[0])
        sizes = [(n - 1) * r * s * s
                 for n, r, s in zip(self.item_nodes, self.item_R, self.item_S)]
        dP = np.zeros(int(np.sum(sizes)))
        dpi = np.zeros(int(np.sum(self.item_S)))
        check(_lib.load_library().mcp_batch_logl_grad_model(self._h, dptr(b), dptr(ll), dptr(g),
                                                            dptr(dP), dptr(dpi)))
        off = np.concatenate([[0], np.cumsum(sizes)])
        soff = np.concatenate([[0], np.cumsum(self.item_S)])
        dPs = [dP[off[i]:off[i + 1]].reshape(self.item_nodes[i] - 1, self.item_R[i],
                                             self.item_S[i], self.item_S[i])
               for i in range(self.n_items)]
        return ll, g, dPs, [dpi[soff[i]:soff[i + 1]] for i in range(self.n_items)]

    def split(self, grad_concat):
        return [grad_concat[self.blv_off[i]:self.blv_off[i + 1]] for i in range(self.n_items)]
