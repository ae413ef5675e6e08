"""NUTS over branch lengths with refraction at zero, driven by the device
leapfrog (mcp_hmc_leapfrog), and the reference's ESS estimator.

The sampler restates the tree sampler of MCPhylo for a fixed topology:
  nuts_sub! / buildtree (src/samplers/pnuts.jl:179-419): multinomial
      trajectory sampling with biased progressive sampling at the top level
      and uniform progressive sampling inside subtrees, divergence check
      `lu + (logp' - logp0) < 1000`, acceptance statistic of the last
      doubling for step-size adaptation;
  refraction! (src/samplers/tree_hamiltonian/refraction.jl:2-91): the
      leapfrog with time-ordered reflection at zero, molifier and scale_fac
      (auxilliary.jl:113-119) -- executed on the device, without the NNI
      topology attempt; with it, ref_NNI! (refraction.jl:35-91) in
      NNIRefractionLeapfrog (host loop, logL from the engine);
  nutsepsilon (auxilliary.jl:6-51): initial step size;
  dual_averaging / setadapt! (pnuts.jl:145-176): step-size adaptation with
      (gamma, kappa, t0) = (0.05, 0.75, 10) and mu = log(10 eps);
  ESS = min((SD / MCSE_bm)^2, n), batch size 100 (src/output/stats.jl:167-172,
      src/output/mcse.jl:26-34).
Differences, each deliberate: the U-turn test is the Euclidean criterion on
branch-length vectors (the reference's `proj_euc`, pnuts.jl:454-481, which
equals its BHV test for a fixed topology); the two trajectory ends integrate
in directions -1 and +1 (pnuts.jl:197-198 gives both -1); a NaN acceptance
statistic counts as 0 (the reference adds -Inf).

Chains are Python generators that yield leapfrog requests (x, r, g, eps,
delta) and receive the new state; `run_nuts` advances any number of chains
in lockstep, one batched device leapfrog per step, so the chains of a
ragged batch share every launch.
"""

from __future__ import annotations

import math
import time

import numpy as np

__all__ = ["HState", "NUTSChain", "run_nuts", "ess_bm", "mcse_bm", "EngineLeapfrog",
           "HostLeapfrog", "SurrogateLeapfrog", "run_da_nuts", "exponential_logprior",
           "NativeNUTS", "run_da_nuts_native", "TreePosition", "nni_parent",
           "NNIRefractionLeapfrog"]


class HState:
    """Phase-space point: position x, momentum r, gradient g, log posterior lf."""

    __slots__ = ("x", "r", "g", "lf")

    def __init__(self, x, r, g, lf):
        self.x, self.r, self.g, self.lf = x, r, g, lf

    def copy(self):
        return HState(self.x.copy(), self.r.copy(), self.g.copy(), self.lf)

    def neg_energy(self) -> float:
        """hamiltonian() of the reference: lf - r.r/2 (states.jl:56)."""
        return self.lf - 0.5 * float(self.r @ self.r)


def _logaddexp(a, b):
    if a == -math.inf:
        return b
    if b == -math.inf:
        return a
    m = max(a, b)
    return m + math.log(math.exp(a - m) + math.exp(b - m))


def _no_uturn(minus: HState, plus: HState) -> bool:
    dx = plus.x - minus.x
    return float(dx @ minus.r) >= 0.0 and float(dx @ plus.r) >= 0.0


class _Meta:
    __slots__ = ("alpha", "nalpha", "leapfrogs")

    def __init__(self):
        self.alpha = 0.0
        self.nalpha = 0
        self.leapfrogs = 0


class NUTSChain:
    """One chain: current position, step size and its adaptation state."""

    def __init__(self, x0, rng, target: float = 0.6, tree_depth: int = 10,
                 epsilon: float | None = None, jitter: float = 0.0, molifier: bool = True):
        """molifier=True keeps the reference's smoothing of positions below
        delta = 2 eps (auxilliary.jl:113-119); its zero-length step at the
        start of every transition moves such positions, so the transition
        is not reversible there.  molifier=False sets delta to 1e-300: pure
        reflection at zero, an exact sampler of the posterior."""
        self.molifier = bool(molifier)
        # subok: a TreePosition keeps its topology
        self.x = np.array(x0, dtype=np.float64, subok=True)
        self.topologies: list[np.ndarray] = []  # per draw, for TreePosition chains
        self.rng = rng
        self.target = float(target)
        self.tree_depth = int(tree_depth)
        self.epsilon = epsilon
        self.jitter = float(jitter)
        # NUTSstepadapter: m, s_bar, x_bar; params mu, delta=target, gamma, kappa, t0
        self.m = 0
        self.s_bar = 0.0
        self.x_bar = 0.0
        self.mu = 0.5
        self.gamma, self.kappa, self.t0 = 0.05, 0.75, 10
        self.metro_acc_prob = 0.0
        self.adapting = False
        self.lf = math.nan
        self.depths: list[int] = []
        self.leapfrogs = 0

    # -- requests -----------------------------------------------------------
    def _step(self, s: HState, eps: float, delta: float):
        """Yield one leapfrog request; returns the new state."""
        if not self.molifier:
            delta = 1e-300
        res = yield (s.x.copy(), s.r.copy(), s.g.copy(), eps, delta)
        self.leapfrogs += 1
        return res

    def _find_epsilon(self):
        """nutsepsilon (auxilliary.jl:6-51) with a cap on the search."""
        n = self.x.shape[0]
        log_target = math.log(self.target)
        zero = np.zeros(n)
        s = yield from self._step(HState(self.x, self.rng.standard_normal(n), zero, 0.0), 0.0,
                                  2.0)
        g0, lf0 = s.g.copy(), s.lf
        eps = 1.0
        x1 = HState(s.x.copy(), self.rng.standard_normal(n), g0, lf0)
        H0 = -x1.neg_energy()
        s2 = yield from self._step(x1.copy(), eps, 2 * eps)
        prob = H0 + s2.neg_energy()
        direction = 1 if prob > log_target else -1
        for _ in range(60):
            if not (prob > log_target if direction == 1 else prob < log_target):
                break
            eps = 2 * eps if direction == 1 else 0.5 * eps
            x1.r = self.rng.standard_normal(n)
            H0 = -x1.neg_energy()
            s2 = yield from self._step(x1.copy(), eps, 2 * eps)
            p = H0 + s2.neg_energy()
            prob = p if math.isfinite(p) else -math.inf
        return eps

    def _buildtree(self, s: HState, direction: int, j: int, eps: float, delta: float,
                   logp0: float, lu: float, meta: _Meta):
        """Subtree of 2^j leapfrogs from s in `direction` (buildtree,
        pnuts.jl:300-419).  Returns (xprime, far, near, n', s', log_sum_weight):
        the sample drawn from the subtree, its end away from s, its end next
        to s."""
        if j == 0:
            s1 = yield from self._step(s, direction * eps, delta)
            logpprime = -s1.neg_energy()
            lsw = logp0 - logpprime
            if not math.isfinite(lsw):
                lsw = -math.inf
            sprime = (lu + (logpprime - logp0)) < 1000.0
            a = math.exp(min(0.0, logp0 - logpprime)) if math.isfinite(logpprime) else 0.0
            meta.alpha += a
            meta.nalpha += 1
            nprime = 1 if (logp0 + lu) < logpprime else 0
            return s1, s1, s1, nprime, bool(sprime), lsw
        xprime, far, near, n1, s1, lsw_init = yield from self._buildtree(
            s, direction, j - 1, eps, delta, logp0, lu, meta)
        if not s1:
            return xprime, far, near, n1, s1, lsw_init
        x2, far2, _, n2, s2, lsw_final = yield from self._buildtree(
            far, direction, j - 1, eps, delta, logp0, lu, meta)
        ls_final = _logaddexp(lsw_init, lsw_final)
        if lsw_final > ls_final or (math.isfinite(lsw_final) and
                                    self.rng.random() < math.exp(lsw_final - ls_final)):
            xprime = x2
        ok = _no_uturn(near, far2) if direction == 1 else _no_uturn(far2, near)
        return xprime, far2, near, n1 + n2, bool(s2 and ok), ls_final

    def iteration(self, adapt: bool):
        """One NUTS transition (nuts_sub!, pnuts.jl:179-295) plus step-size
        adaptation (sample!, pnuts.jl:103-143).  Generator."""
        if self.epsilon is None or not math.isfinite(self.epsilon):
            self.epsilon = yield from self._find_epsilon()
        if adapt and not self.adapting:
            self.m = 0
            self.mu = math.log(10.0 * self.epsilon)
        self.adapting = adapt
        if adapt:
            self.m += 1
            eps = self.epsilon
        else:
            if self.m > 0:
                self.epsilon = math.exp(self.x_bar)
            eps = self.epsilon * (1.0 + self.jitter * (2.0 * self.rng.random() - 1.0))
        delta = 2.0 * self.epsilon
        n = self.x.shape[0]
        r = self.rng.standard_normal(n)
        s0 = yield from self._step(HState(self.x, r, np.zeros(n), 0.0), 0.0, delta)
        minus, plus, chosen = s0.copy(), s0.copy(), s0
        logp0 = -s0.neg_energy()
        lu = math.log(self.rng.random())
        log_sum_weight = 0.0
        meta = _Meta()
        j = 0
        while j < self.tree_depth:
            direction = -1 if self.rng.random() > 0.5 else 1
            meta.alpha, meta.nalpha = 0.0, 0
            if direction == -1:
                xprime, minus, _, nprime, sprime, lsw = yield from self._buildtree(
                    minus, -1, j, eps, delta, logp0, lu, meta)
            else:
                xprime, plus, _, nprime, sprime, lsw = yield from self._buildtree(
                    plus, 1, j, eps, delta, logp0, lu, meta)
            self.metro_acc_prob = min(1.0, meta.alpha / meta.nalpha) if meta.nalpha else 0.0
            if not sprime:
                break
            if lsw > log_sum_weight or (math.isfinite(lsw) and
                                        self.rng.random() < math.exp(lsw - log_sum_weight)):
                chosen = xprime
            log_sum_weight = _logaddexp(log_sum_weight, lsw)
            j += 1
            if not _no_uturn(minus, plus):
                break
        self.x = chosen.x.copy()
        self.lf = chosen.lf
        self.depths.append(j)
        if adapt:
            # dual_averaging (pnuts.jl:145-159)
            HT = self.target - self.metro_acc_prob
            eta = 1.0 / (self.m + self.t0)
            self.s_bar = (1.0 - eta) * self.s_bar + eta * HT
            xx = self.mu - self.s_bar * math.sqrt(self.m) / self.gamma
            xe = self.m ** (-self.kappa)
            self.x_bar = (1.0 - xe) * self.x_bar + xe * xx
            self.epsilon = math.exp(xx)
        return self.x


class EngineLeapfrog:
    """Leapfrog provider backed by a device engine (Plan or Batch): item i is
    chain i.  With `record` a list, every step of nonzero length appends
    (chain, x, logL(x)) -- the (state, logL) pairs a surrogate trains on."""

    def __init__(self, engine, n_items: int, n_branches, record=None):
        self.engine = engine
        self.n_items = n_items
        self.nb = np.asarray(n_branches)
        self.off = np.concatenate([[0], np.cumsum(self.nb)])
        T = int(self.off[-1])
        self.x = np.zeros(T)
        self.r = np.zeros(T)
        self.g = np.zeros(T)
        self.ed = np.zeros(2 * n_items)
        self.record = record

    def __call__(self, reqs):
        for i, (x, r, g, eps, delta) in enumerate(reqs):
            a, b = self.off[i], self.off[i + 1]
            self.x[a:b] = x
            self.r[a:b] = r
            self.g[a:b] = g
            self.ed[2 * i] = eps
            self.ed[2 * i + 1] = delta
        lp, ll = self.engine.leapfrog(self.ed, self.x, self.r, self.g)
        if self.record is not None:
            for i in range(self.n_items):
                if self.ed[2 * i] != 0.0:
                    self.record.append((i, self.x[self.off[i]:self.off[i + 1]].copy(),
                                        float(ll[i])))
        return [HState(self.x[self.off[i]:self.off[i + 1]].copy(),
                       self.r[self.off[i]:self.off[i + 1]].copy(),
                       self.g[self.off[i]:self.off[i + 1]].copy(), float(lp[i]))
                for i in range(self.n_items)]


class HostLeapfrog:
    """Leapfrog provider from a host function logfgrad(x) -> (lf, grad)
    (the CPU path; also used by the tests with toy targets)."""

    def __init__(self, logfgrad, prior_scale: float | None = None, prior=None):
        """prior_scale: exponential branch-length prior; prior: any callable
        x -> (logprior, gradient) instead."""
        self.logfgrad = logfgrad
        self.prior_scale = prior_scale
        self.prior = prior

    def one(self, x, r, g, eps, delta):
        from_oracle = _host_refraction(x, r, g, eps, delta)
        xm, r1 = from_oracle
        lf, gr = self.logfgrad(xm)
        if self.prior is not None:
            lpr, gpr = self.prior(xm)
            lf += lpr
            gr = gr + gpr
        elif self.prior_scale is not None:
            lf += float(np.sum(-math.log(self.prior_scale) - xm / self.prior_scale))
            gr = gr - 1.0 / self.prior_scale
        gr = gr * np.where(xm < delta, xm / delta, 1.0)
        r1 = r1 + (eps * 0.5) * gr
        return HState(xm, r1, gr, lf)

    def __call__(self, reqs):
        return [self.one(*q) for q in reqs]


class SurrogateLeapfrog:
    """Leapfrog provider on the surrogate posterior logL~(x) + logprior(x)
    (mcp_mlp_leapfrog): chain i is row i."""

    def __init__(self, surrogate, n_chains: int, prior_scale: float | None = None):
        self.s = surrogate
        self.n = int(n_chains)
        self.prior_scale = prior_scale
        d = surrogate.d
        self.x = np.zeros((self.n, d))
        self.r = np.zeros((self.n, d))
        self.g = np.zeros((self.n, d))
        self.ed = np.zeros(2 * self.n)

    def __call__(self, reqs):
        for i, (x, r, g, eps, delta) in enumerate(reqs):
            self.x[i] = x
            self.r[i] = r
            self.g[i] = g
            self.ed[2 * i] = eps
            self.ed[2 * i + 1] = delta
        lp = self.s.leapfrog(self.ed, self.x, self.r, self.g, self.prior_scale)
        return [HState(self.x[i].copy(), self.r[i].copy(), self.g[i].copy(), float(lp[i]))
                for i in range(self.n)]


def _host_refraction(x, r, g, eps, delta, max_reflect=None):
    """First half of refraction! on the host (same steps as the device)."""
    r = r + (eps * 0.5) * g
    b = x.copy()
    t = 0.0
    cap = 64 * max(1, x.shape[0]) if max_reflect is None else max_reflect
    for _ in range(cap):
        tb = b + (eps - t) * r
        if not np.any(tb <= 0.0):
            break
        with np.errstate(divide="ignore", invalid="ignore"):
            tl = tb / np.abs(r)
        k = int(np.argmin(tl))
        temp = eps - t + tl[k]
        b = np.abs(b + temp * r)
        r = r.copy()
        r[k] = -r[k]
        t = eps + tl[k]
    y = np.abs(b + (eps - t) * r)
    xm = np.where(y >= delta, y, (y * y + delta * delta) / (2.0 * delta))
    return xm, r


# ---------------------------------------------------------------------------
# refraction with NNI moves (refraction.jl:2-91)
# ---------------------------------------------------------------------------

class TreePosition(np.ndarray):
    """Branch lengths by node num (blv) that carry their topology: `parent`
    (int32 [N], 1-based mother num, 0 at the root num N).  Copies and
    arithmetic keep the attribute, so a NUTS chain moves tree positions
    through its trajectories unchanged (the reference's state holds the
    tree itself)."""

    def __new__(cls, blv, parent):
        obj = np.array(blv, dtype=np.float64).view(cls)
        obj.parent = np.array(parent, dtype=np.int32)
        return obj

    def __array_finalize__(self, obj):
        self.parent = getattr(obj, "parent", None)


def nni_parent(parent, v: int, pick: int):
    """Nearest-neighbour interchange across the branch above node num v:
    child `pick` (0 or 1, children in ascending num) of v trades places with
    v's sibling.  Returns (new parent array, 1), or (parent, 0) when v is a
    leaf or the root or v or its mother is not bifurcating.  Node nums and
    so the branch-length vector are unchanged.

    The reference calls MCPhyloTree's NNI!(tree, index) (refraction.jl:65),
    a dependency this repository does not hold, so the choice of swap is a
    restatement (parity unpinned): one of the two interchanges around the
    branch, picked by the caller."""
    par = np.asarray(parent, dtype=np.int32)
    N = par.shape[0]
    if v < 1 or v >= N:
        return par, 0
    kids = np.flatnonzero(par == v) + 1
    p = int(par[v - 1])
    sibs = [int(u) for u in np.flatnonzero(par == p) + 1 if u != v]
    if kids.shape[0] != 2 or len(sibs) != 1:
        return par, 0
    c = int(kids[pick & 1])
    new = par.copy()
    new[c - 1] = p
    new[sibs[0] - 1] = v
    return new, 1


def _molify(y, delta):
    return np.where(y >= delta, y, (y * y + delta * delta) / (2.0 * delta))


class NNIRefractionLeapfrog:
    """refraction! with NNI attempts (refraction.jl:2-91) on one engine.

    Positions are TreePosition.  A branch whose end is internal and that
    reaches zero during the step triggers ref_NNI!'s attempt: the log
    posterior of the current tree (at the molified reflection point) and of
    the interchanged tree are compared, and the interchange is kept when the
    branch's squared momentum exceeds 2 (U_before - U_after), the momentum
    then shrinking to sqrt(r^2 - 2 dU).  logL comes from the engine (its
    topology is switched with set_topology); prior(x) -> (logprior, grad)
    adds the tree prior.  Counts nni (made) and att_nni (attempted) as the
    reference does.  Multiple chains share the engine one request at a
    time."""

    def __init__(self, plan, prior=None, rng=None, max_reflect: int | None = None):
        self.plan = plan
        self.prior = prior
        self.rng = rng if rng is not None else np.random.default_rng(0)
        self.max_reflect = max_reflect
        self.nni = 0
        self.att_nni = 0

    def _use(self, parent):
        if not np.array_equal(self.plan.parent, parent):
            self.plan.set_topology(parent)

    def _logpost(self, x):
        lf = self.plan.logl(np.asarray(x, dtype=np.float64))
        if self.prior is not None:
            lf += self.prior(np.asarray(x))[0]
        return lf

    def one(self, x, r, g, eps, delta):
        topo = np.array(x.parent, dtype=np.int32)
        N = topo.shape[0]
        internal = np.zeros(N - 1, dtype=bool)
        for u in topo[:N - 1]:
            if 1 <= u <= N - 1:
                internal[u - 1] = True
        blv = np.asarray(x, dtype=np.float64).copy()
        r = np.asarray(r, dtype=np.float64) + (eps * 0.5) * np.asarray(g)
        tmp = blv + eps * r
        t = 0.0
        cap = 64 * max(1, blv.shape[0]) if self.max_reflect is None else self.max_reflect
        for _ in range(cap):
            if not np.any(tmp <= 0.0):
                break
            with np.errstate(divide="ignore", invalid="ignore"):
                tl = tmp / np.abs(r)
            k = int(np.argmin(tl))
            temp = eps - t + tl[k]
            blv = np.abs(blv + temp * r)
            r = r.copy()
            r[k] = -r[k]
            if internal[k]:
                xm = _molify(blv, delta)
                self._use(topo)
                u_before = self._logpost(xm)
                cand, made = nni_parent(topo, k + 1, int(self.rng.integers(2)))
                if made:
                    self.att_nni += 1
                    self._use(cand)
                    u_after = self._logpost(xm)
                    d_u = 2.0 * (u_before - u_after)
                    v2 = r[k] * r[k]
                    if v2 > d_u:
                        self.nni += made
                        # positive whatever the direction of integration, as
                        # in the reference (refraction.jl:76)
                        r[k] = math.sqrt(v2 - d_u)
                        topo = cand
            t = eps + tl[k]
            tmp = blv + (eps - t) * r
        xm = _molify(tmp, delta)
        self._use(topo)
        lf, gr = self.plan.logl_grad(xm)
        if self.prior is not None:
            lpr, gpr = self.prior(xm)
            lf += lpr
            gr = gr + gpr
        gr = gr * np.where(xm < delta, xm / delta, 1.0)
        r = r + (eps * 0.5) * gr
        return HState(TreePosition(xm, topo), r, gr, float(lf))

    def __call__(self, reqs):
        return [self.one(*q) for q in reqs]


def run_nuts(chains, leapfrog, n_iter: int, n_adapt: int = 0, trace=None):
    """Advance every chain through n_iter transitions (the first n_adapt with
    step-size adaptation) in lockstep: each round issues one batched
    leapfrog.  Chains that finish early repeat a zero-length step until all
    are done.  Returns (draws [n_chains][n_iter][n_branches], lf trace,
    wall seconds, leapfrog rounds)."""
    n = len(chains)
    draws = [[] for _ in range(n)]
    lfs = [[] for _ in range(n)]
    done_iters = [0] * n
    gens = [None] * n
    reqs = [None] * n

    def start(i):
        gens[i] = chains[i].iteration(done_iters[i] < n_adapt)
        reqs[i] = next(gens[i])

    for i in range(n):
        start(i)
    rounds = 0
    t0 = time.perf_counter()
    while any(g is not None for g in gens):
        live = [q if q is not None else _idle(chains[i]) for i, q in enumerate(reqs)]
        results = leapfrog(live)
        rounds += 1
        for i in range(n):
            if gens[i] is None:
                continue
            try:
                reqs[i] = gens[i].send(results[i])
            except StopIteration:
                if done_iters[i] >= n_adapt:
                    draws[i].append(chains[i].x.copy())
                    lfs[i].append(chains[i].lf)
                    if getattr(chains[i].x, "parent", None) is not None:
                        chains[i].topologies.append(chains[i].x.parent.copy())
                done_iters[i] += 1
                if done_iters[i] < n_iter + n_adapt:
                    start(i)
                else:
                    gens[i] = None
                    reqs[i] = None
                if trace is not None:
                    trace(i, done_iters[i])
    wall = time.perf_counter() - t0
    return [np.array(d) for d in draws], [np.array(v) for v in lfs], wall, rounds


def _lockstep_iteration(chains, leapfrog, adapt: bool) -> int:
    """One NUTS transition of every chain, the chains' leapfrogs batched into
    shared calls; returns the number of calls."""
    gens = [c.iteration(adapt) for c in chains]
    reqs = [next(g) for g in gens]
    rounds = 0
    while any(g is not None for g in gens):
        live = [q if q is not None else _idle(chains[i]) for i, q in enumerate(reqs)]
        results = leapfrog(live)
        rounds += 1
        for i, g in enumerate(gens):
            if g is None:
                continue
            try:
                reqs[i] = g.send(results[i])
            except StopIteration:
                gens[i] = None
                reqs[i] = None
    return rounds


def exponential_logprior(X, scale: float) -> np.ndarray:
    """Sum over branches of the Exponential(mean scale) log density, per row
    (exponentialBL, src/Likelihood/Prior.jl:50-56,80-83)."""
    X = np.atleast_2d(X)
    return np.sum(-math.log(scale) - X / scale, axis=1)


def run_da_nuts(chains, leapfrog, exact_logpost, n_iter: int, n_adapt: int = 0,
                min_accept: float = 0.0, window: int = 50, fallback=None, trace=None,
                record=None, stage1: int = 1):
    """Delayed-acceptance NUTS.  Stage one: a NUTS transition of every chain
    integrated on the surrogate posterior by `leapfrog` (a SurrogateLeapfrog);
    its draw x' is a proposal.  Stage two: one batched exact evaluation
    `exact_logpost(X) -> [n_chains]` and the Metropolis-Hastings correction
        log a = (lp(x') - lp~(x')) - (lp(x) - lp~(x)),
    which leaves the exact posterior invariant because the NUTS transition is
    reversible with respect to the surrogate posterior (so is `stage1` > 1
    repetitions of it, a power of a self-adjoint kernel).  lp~ of the current
    state is the value the surrogate leapfrog returned for it (the first one
    from a zero-length step with the molifier off), so both stages compare
    values from the same kernels.

    Fallback: after adaptation, when the second-stage acceptance over the
    last `window` iterations (all chains) drops below `min_accept`, the
    remaining iterations run plain NUTS with `fallback` (an exact leapfrog
    provider).  With `record` a list, every second-stage evaluation appends
    (chain, x', exact lp(x')) -- the pairs that refine the surrogate where it
    proposes.  Returns (draws, exact lp of the draws, wall seconds, stats)."""
    n = len(chains)
    d = chains[0].x.shape[0]
    t0 = time.perf_counter()
    zero = np.zeros(d)
    first = leapfrog([(c.x.copy(), zero, zero, 0.0, 1e-300) for c in chains])
    sur = np.array([s.lf for s in first])
    lp = np.asarray(exact_logpost(np.stack([c.x for c in chains])), dtype=np.float64).copy()
    draws = [[] for _ in range(n)]
    lps = [[] for _ in range(n)]
    stats = {"rounds": 1, "exact_evals": n, "proposals": 0, "accepted": 0,
             "fallback_at": None, "accept_trace": []}
    recent: list[float] = []
    total = n_adapt + n_iter
    it = 0
    while it < total:
        adapt = it < n_adapt
        x_old = [c.x.copy() for c in chains]
        for _ in range(max(1, int(stage1))):
            stats["rounds"] += _lockstep_iteration(chains, leapfrog, adapt)
        X = np.stack([c.x for c in chains])
        lp_new = np.asarray(exact_logpost(X), dtype=np.float64)
        stats["exact_evals"] += n
        if record is not None:
            record.extend((i, X[i].copy(), float(lp_new[i])) for i in range(n))
        acc = 0
        for i, c in enumerate(chains):
            la = (lp_new[i] - c.lf) - (lp[i] - sur[i])
            if la >= 0.0 or (math.isfinite(la) and c.rng.random() < math.exp(la)):
                lp[i] = lp_new[i]
                sur[i] = c.lf
                acc += 1
            else:
                c.x = x_old[i]
                c.lf = sur[i]
        stats["proposals"] += n
        stats["accepted"] += acc
        stats["accept_trace"].append(acc / n)
        if not adapt:
            for i, c in enumerate(chains):
                draws[i].append(c.x.copy())
                lps[i].append(lp[i])
            recent.append(acc / n)
            recent = recent[-window:]
        it += 1
        if trace is not None:
            trace(it)
        if (fallback is not None and not adapt and len(recent) == window
                and float(np.mean(recent)) < min_accept):
            stats["fallback_at"] = it
            rest, rest_lf, _, rounds = run_nuts(chains, fallback, total - it, 0)
            stats["rounds"] += rounds
            for i in range(n):
                draws[i].extend(list(rest[i]))
                lps[i].extend(list(rest_lf[i]))
            break
    wall = time.perf_counter() - t0
    stats["accept_rate"] = stats["accepted"] / max(1, stats["proposals"])
    return [np.array(v) for v in draws], [np.array(v) for v in lps], wall, stats


def _idle(chain):
    z = np.zeros_like(chain.x)
    return (chain.x.copy(), z, z, 0.0, 2.0 * (chain.epsilon or 1.0))


def mcse_bm(x, size: int = 100) -> float:
    """Batch-means MCSE (src/output/mcse.jl:26-34): sem of the batch means."""
    x = np.asarray(x, dtype=np.float64)
    m = x.shape[0] // size
    if m < 2:
        raise ValueError(f"iterations are < {2 * size} and batch size is > {x.shape[0] // 2}")
    mbar = x[: m * size].reshape(m, size).mean(axis=1)
    return float(np.std(mbar, ddof=1) / math.sqrt(m))


def ess_bm(x, size: int = 100) -> float:
    """ESS = min((SD / MCSE_bm)^2, n) (src/output/stats.jl:167-172)."""
    x = np.asarray(x, dtype=np.float64)
    sd = float(np.std(x, ddof=1))
    se = mcse_bm(x, size)
    if se == 0.0:
        return float(x.shape[0])
    return float(min((sd / se) ** 2, x.shape[0]))


class NativeNUTS:
    """The native lockstep NUTS driver (mcp_nuts_*): the same sampler as
    NUTSChain/run_nuts, with every chain's trajectory building in C++ and one
    batched leapfrog per round.  Chains may have different numbers of
    branches (the items of a ragged batch)."""

    def __init__(self, x0, seeds, target: float = 0.8, tree_depth: int = 10,
                 molifier: bool = True):
        import ctypes
        from . import _lib
        self._lib = _lib
        self.dims = np.array([len(x) for x in x0], dtype=np.int32)
        self.off = np.concatenate([[0], np.cumsum(self.dims)]).astype(np.int64)
        x = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64) for v in x0]))
        sd = np.ascontiguousarray(seeds, dtype=np.uint64)
        if sd.shape[0] != self.dims.shape[0]:
            raise ValueError("one seed per chain")
        self._h = ctypes.c_void_p()
        _lib.check(_lib.load_library().mcp_nuts_create(
            ctypes.byref(self._h), len(self.dims), _lib.iptr(self.dims), _lib.dptr(x),
            sd.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), float(target), int(tree_depth),
            1 if molifier else 0))
        self._keep = None
        self.n = len(self.dims)
        self.total = int(self.off[-1])

    def close(self):
        if self._h and self._h.value:
            self._lib.load_library().mcp_nuts_destroy(self._h)
            import ctypes
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def use_engine(self, engine) -> None:
        self._keep = engine
        self._lib.check(self._lib.load_library().mcp_nuts_use_batch(self._h, engine.handle))

    def use_surrogate(self, surrogate, prior_scale: float | None = None) -> None:
        self._keep = surrogate
        self._lib.check(self._lib.load_library().mcp_nuts_use_mlp(
            self._h, surrogate._h, -1.0 if prior_scale is None else float(prior_scale)))

    def use_callback(self, leapfrog) -> None:
        """leapfrog(eps_delta [2n], x, r, g [total], logpost [n]) updates the
        arrays in place (numpy views of the driver's buffers)."""
        total, n = self.total, self.n

        def cb(ctx, nch, ed, x, r, g, lp):
            try:
                leapfrog(np.ctypeslib.as_array(ed, (2 * n,)), np.ctypeslib.as_array(x, (total,)),
                         np.ctypeslib.as_array(r, (total,)), np.ctypeslib.as_array(g, (total,)),
                         np.ctypeslib.as_array(lp, (n,)))
                return 0
            except Exception:
                return 1
        self._keep = self._lib.LEAPFROG_FN(cb)
        self._lib.check(self._lib.load_library().mcp_nuts_use_callback(self._h, self._keep, None))

    def run(self, n_iter: int, adapt: bool = False, keep_draws: bool = True):
        """n_iter transitions of every chain.  Returns (draws [chain] ->
        [n_iter, dim], logpost [n_iter, n_chains], wall seconds, rounds,
        leapfrogs)."""
        import ctypes
        d = np.zeros((n_iter, self.total)) if keep_draws else None
        lp = np.zeros((n_iter, self.n))
        st = np.zeros(2, dtype=np.int64)
        t0 = time.perf_counter()
        self._lib.check(self._lib.load_library().mcp_nuts_run(
            self._h, int(n_iter), 1 if adapt else 0, self._lib.dptr(d), self._lib.dptr(lp),
            st.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        wall = time.perf_counter() - t0
        draws = None if d is None else [d[:, self.off[i]:self.off[i + 1]] for i in range(self.n)]
        return draws, lp, wall, int(st[0]), int(st[1])

    def state(self):
        """(x concatenated, step sizes, log posteriors, last tree depths)."""
        x = np.zeros(self.total)
        eps, lf = np.zeros(self.n), np.zeros(self.n)
        dep = np.zeros(self.n, dtype=np.int32)
        self._lib.check(self._lib.load_library().mcp_nuts_get(
            self._h, self._lib.dptr(x), self._lib.dptr(eps), self._lib.dptr(lf),
            self._lib.iptr(dep)))
        return x, eps, lf, dep

    def set_state(self, x=None, eps=None, lf=None) -> None:
        def arr(a):
            return None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        x, eps, lf = arr(x), arr(eps), arr(lf)
        self._lib.check(self._lib.load_library().mcp_nuts_set(
            self._h, self._lib.dptr(x), self._lib.dptr(eps), self._lib.dptr(lf)))


def run_da_nuts_native(nuts: NativeNUTS, surrogate, prior_scale, exact_logpost, n_iter: int,
                       n_adapt: int = 0, trace=None):
    """Delayed-acceptance NUTS with the native driver as stage one (on the
    surrogate posterior, `nuts` set up with use_surrogate) and one batched
    exact evaluation per iteration as stage two; the acceptance is that of
    run_da_nuts.  Returns (draws [chain] -> [n_iter, dim], exact lp
    [n_iter, n_chains], wall seconds, stats)."""
    n, d = nuts.n, int(nuts.dims[0])
    t0 = time.perf_counter()
    x, _, _, _ = nuts.state()
    X = x.reshape(n, d)
    ed = np.tile([0.0, 1e-300], n)
    xx, rr, gg = X.copy(), np.zeros((n, d)), np.zeros((n, d))
    sur = surrogate.leapfrog(ed, xx, rr, gg, prior_scale)
    lp = np.asarray(exact_logpost(X), dtype=np.float64).copy()
    draws = np.zeros((n_iter, n, d))
    lps = np.zeros((n_iter, n))
    stats = {"exact_evals": n, "proposals": 0, "accepted": 0, "rounds": 1}
    rng = np.random.default_rng(int(nuts.dims.sum()) + 7)
    for it in range(n_adapt + n_iter):
        x_old = X.copy()
        _, _, _, rounds, _ = nuts.run(1, adapt=it < n_adapt, keep_draws=False)
        stats["rounds"] += rounds
        xn, _, sur_new, _ = nuts.state()
        Xn = xn.reshape(n, d)
        lp_new = np.asarray(exact_logpost(Xn), dtype=np.float64)
        stats["exact_evals"] += n
        la = (lp_new - sur_new) - (lp - sur)
        acc = (la >= 0.0) | (np.isfinite(la) & (rng.random(n) < np.exp(np.minimum(la, 0.0))))
        X = np.where(acc[:, None], Xn, x_old)
        lp = np.where(acc, lp_new, lp)
        sur = np.where(acc, sur_new, sur)
        if not acc.all():
            nuts.set_state(x=X.reshape(-1), lf=sur)
        stats["proposals"] += n
        stats["accepted"] += int(acc.sum())
        if it >= n_adapt:
            draws[it - n_adapt] = X
            lps[it - n_adapt] = lp
        if trace is not None:
            trace(it + 1)
    wall = time.perf_counter() - t0
    stats["accept_rate"] = stats["accepted"] / max(1, stats["proposals"])
    return [draws[:, i, :] for i in range(n)], lps, wall, stats
