#!/usr/bin/env python3
"""Benchmark: logL + branch-length gradient site-pattern evaluations per
second on one MI355X.

Workload (default): 100-taxon random rooted binary tree, branch lengths
~ Exp(mean 0.05), 4-state GTR (exchangeabilities ~ Dirichlet(1), base
frequencies ~ Dirichlet(5)), 4 discrete-gamma categories (alpha = 0.5),
10,000 distinct simulated site patterns with integer weights; FP64.

One step = one full evaluation: upload of the branch lengths, P/dP build,
every post-order and pre-order level, reductions and download of
(logL, grad[198]).  Inputs (tip codes, model) are resident in HBM before the
timed region.  value = patterns * steps / wall seconds.

Prints one JSON line on stdout; progress goes to stderr.
"""

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def load_package():
    if "mcphylo_amd" in sys.modules:
        return sys.modules["mcphylo_amd"]
    pkg = os.path.join(ROOT, "mcphylo.jl_amd")
    spec = importlib.util.spec_from_file_location(
        "mcphylo_amd", os.path.join(pkg, "__init__.py"), submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mcphylo_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


# loaded at import so that spawned workers can unpickle the generators
mcp = load_package()


def make_batch_data(n_trees, workers):
    """Ragged batch inputs, generated in worker processes started with
    'spawn' before this process touches the GPU (workers = 0: in-process)."""
    if workers <= 0:
        return mcp.synthetic.make_batched_workload(n_trees=n_trees)
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    if mcp._lib._lib is not None:
        raise RuntimeError("batch data must be generated before the HIP library is loaded")
    with ProcessPoolExecutor(workers, mp_context=mp.get_context("spawn")) as pool:
        return mcp.synthetic.make_batched_workload(n_trees=n_trees, pool=pool)


def bench_batched(data, steps, warmup, sample_trees, schedule):
    """Batched logL+grad over all (tree, chain) items in one engine, and the
    plain single-tree launch sequence on a sample of the same trees."""
    datasets, items, models, blv = data
    b = mcp.Batch(datasets, items)
    b.set_schedule("levels" if schedule == "auto" else schedule)
    for i, m in enumerate(models):
        b.set_model(i, mcp.Model(*m))
    b.logl_grad(blv)
    tiles, tune_rows = b.autotune(reps=2)
    log(f"  autotuned tiles (columns per thread by state count): {tiles}")
    sched_rows = []
    if schedule == "auto":
        _, sched_rows = b.autotune_schedule(reps=2)
        tiles, tune_rows = b.autotune(reps=2)
        _, sched_rows = b.autotune_schedule(reps=2)
    sched, hwidth = b.schedule()
    log(f"  schedule: {sched}" + (f" (hybrid width {hwidth})" if sched == "hybrid" else ""))
    n_items, total_patterns = b.sizes()
    for _ in range(warmup):
        b.logl_grad(blv)
    t0 = time.perf_counter()
    for _ in range(steps):
        b.logl_grad(blv)
    wall = time.perf_counter() - t0
    b.set_timing(True)
    b.logl_grad(blv)
    ms, ph = b.launch_times()
    lb = b.launch_bytes()
    b.set_timing(False)
    bytes_fwd, bytes_adj = b.algorithmic_bytes()
    b.close()
    fwd_ms, adj_ms = float(np.sum(ms[ph == 2])), float(np.sum(ms[ph == 3]))
    out = {"items": n_items, "patterns_per_step": total_patterns,
           "schedule": sched, "hybrid_width": hwidth if sched == "hybrid" else None,
           "autotuned_cols_per_thread": {str(k): v for k, v in tiles.items()},
           "autotune_ms": [list(r) for r in tune_rows],
           "autotune_schedule_ms": [list(r) for r in sched_rows],
           "ms_per_step": 1e3 * wall / steps,
           "value": total_patterns * steps / wall,
           "algorithmic_GB_per_step": (bytes_fwd + bytes_adj) / 1e9,
           "launches": int(ms.shape[0])}
    if fwd_ms > 0:
        out["post_order_ms"] = fwd_ms
        out["post_order_GBs"] = float(np.sum(lb[ph == 2])) / (fwd_ms * 1e-3) / 1e9
    if adj_ms > 0:
        out["pre_order_ms"] = adj_ms
        out["pre_order_GBs"] = float(np.sum(lb[ph == 3])) / (adj_ms * 1e-3) / 1e9
    # plain single-tree sequence (one launch per tree level) on the first
    # chain of the first trees
    pat, tsum = 0, 0.0
    offs = np.concatenate([[0], np.cumsum([it["parent"].shape[0] - 1 for it in items])])
    chains = len(items) // len(datasets)
    for t in range(min(sample_trees, len(datasets))):
        i = t * chains
        d = datasets[t]
        L, C = d["tip_codes"].shape
        pb = mcp.Batch([d], [dict(items[i], dataset=0)])
        pb.set_schedule("levels")
        pb.set_model(0, mcp.Model(*models[i]))
        bl = blv[offs[i]:offs[i + 1]]
        pb.logl_grad(bl)
        t1 = time.perf_counter()
        for _ in range(3):
            pb.logl_grad(bl)
        tsum += (time.perf_counter() - t1) / 3
        pat += C
        pb.close()
    out["plain_value"] = pat / tsum
    out["plain_sample_trees"] = int(min(sample_trees, len(datasets)))
    out["batched_over_plain"] = out["value"] / out["plain_value"]
    return out


def nuts_run(chain, leapfrog, n_adapt, n_iter):
    """Adaptation, then n_iter timed NUTS transitions with the host driver;
    ESS (batch means, batch size 100) of tree length and log posterior per
    sampling second."""
    S = mcp.sampler
    S.run_nuts([chain], leapfrog, 0, n_adapt=n_adapt)
    log(f"  adapted: step size {chain.epsilon:.3g}")
    draws, lfs, wall, rounds = S.run_nuts([chain], leapfrog, n_iter, n_adapt=0)
    log(f"  {n_iter} iterations in {wall:.1f} s")
    tl = draws[0].sum(axis=1)
    ess_tl, ess_lf = S.ess_bm(tl), S.ess_bm(lfs[0])
    return {"iterations": n_iter, "wall_s": wall, "leapfrogs": rounds,
            "ess_tree_length": ess_tl, "ess_logpost": ess_lf,
            "ess_per_s": min(ess_tl, ess_lf) / wall,
            "mean_tree_length": float(tl.mean()), "sd_tree_length": float(tl.std(ddof=1)),
            "mean_logpost": float(np.mean(lfs[0])), "step_size": chain.epsilon,
            "mean_tree_depth": float(np.mean(chain.depths[-n_iter:]))}


def native_run(nuts, n_adapt, n_iter):
    """Adaptation, then n_iter timed transitions of every chain with the native
    driver; per chain ESS (batch size 100) of tree length and log posterior,
    summed over chains, per sampling second."""
    S = mcp.sampler
    nuts.run(n_adapt, adapt=True, keep_draws=False)
    draws, lp, wall, rounds, lfs = nuts.run(n_iter)
    ess = [min(S.ess_bm(d.sum(axis=1)), S.ess_bm(lp[:, i])) for i, d in enumerate(draws)]
    tl = np.concatenate([d.sum(axis=1) for d in draws])
    _, eps, _, _ = nuts.state()
    return {"chains": nuts.n, "iterations": n_iter, "wall_s": wall, "leapfrog_rounds": rounds,
            "leapfrogs": lfs, "ess_sum": float(np.sum(ess)), "ess_min_chain": float(np.min(ess)),
            "ess_per_s": float(np.sum(ess)) / wall, "mean_tree_length": float(tl.mean()),
            "mean_step_size": float(np.mean(eps)), "ms_per_round": 1e3 * wall / max(rounds, 1)}


def bench_nuts(n_adapt, n_iter, cpu: bool, batch_chains: int = 64):
    """NUTS on the 20-taxon restriction-site workload: the native driver with
    the device leapfrog (one HIP graph per round) for one chain and for
    `batch_chains` chains sharing every launch; the baseline is the same
    sampler on one chain with the CPU oracle's logL + gradient."""
    S = mcp.sampler
    w = mcp.synthetic.make_workload("restriction20")
    blv = mcp.get_branchlength_vector(w.tree)
    leaf = w.ta.leaf_nums
    out = {"workload": "20-taxon Yule tree, restriction-site model pi=(.5,.5), 1000 simulated "
                       "sites, R=1, exponential(0.1) branch-length prior, NUTS target 0.8, "
                       "tree depth 5 (the reference's PNUTS test setting); ESS = min over tree "
                       "length and log posterior per chain, summed over chains"}
    # a chain's ESS from the reference estimator (batches of 100 draws) needs
    # many batches: the GPU chains run 10x the iterations (20 batches each),
    # the CPU baseline 2x
    for key, nch, iters in (("gpu", 1, 10 * n_iter), ("gpu_batched", batch_chains, 10 * n_iter)):
        b = mcp.Batch([{"x_leaves": w.x[:, :, leaf - 1], "leaf_nums": leaf}],
                      [{"dataset": 0, "parent": w.ta.parent, "n_rates": 1}] * nch)
        for i in range(nch):
            b.set_model(i, mcp.Model(w.U, w.D, w.Uinv, w.mu, w.rates, w.pi))
            b.set_prior_exponential(i, 0.1)
        b.use_graph(True)
        nuts = S.NativeNUTS([blv] * nch, [20261202 + i for i in range(nch)], target=0.8,
                            tree_depth=5)
        nuts.use_engine(b)
        out[key] = native_run(nuts, n_adapt, iters)
        log(f"  {key}: {out[key]['ess_per_s']:.1f} ESS/s over {nch} chain(s)")
        nuts.close()
        b.close()
    out["batched_over_single_ess_per_s"] = out["gpu_batched"]["ess_per_s"] / out["gpu"]["ess_per_s"]
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        oracle.set_threads(min(os.cpu_count() or 1, 16))

        def logfgrad(x):
            return oracle.logl_grad(w.ta, w.x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, x)
        c2 = S.NUTSChain(blv, np.random.default_rng(20261202), target=0.8, tree_depth=5)
        out["cpu"] = nuts_run(c2, S.HostLeapfrog(logfgrad, prior_scale=0.1), n_adapt, 2 * n_iter)
        out["cpu"]["threads"] = min(os.cpu_count() or 1, 16)
        out["gpu_over_cpu_ess_per_s"] = out["gpu"]["ess_per_s"] / out["cpu"]["ess_per_s"]
        out["gpu_batched_over_cpu_ess_per_s"] = (out["gpu_batched"]["ess_per_s"]
                                                 / out["cpu"]["ess_per_s"])
    return out


def bench_da(n_warm, n_iter, n_chains=8, prior=0.1, depth=8, grad_weight=1e4, epochs=30):
    """Delayed-acceptance NUTS on the 200-taxon GTR workload.  A shared warm-up
    (plain NUTS, exact device leapfrog, molifier off) records every leapfrog's
    (x, logL, dlogL/dx) after its first third; plain NUTS and surrogate DA-NUTS
    then continue from the warm-up's states and step sizes for n_iter
    iterations each.  ESS by batch means (batch 25) over the pooled chains."""
    S = mcp.sampler
    w = mcp.synthetic.make_workload("gtr200")
    ta = w.ta
    d = ta.parent.shape[0] - 1
    model = mcp.Model(w.U, w.D, w.Uinv, w.mu, w.rates, w.pi)
    b = mcp.Batch([{"x_leaves": w.x[:, :, ta.leaf_nums - 1], "leaf_nums": ta.leaf_nums,
                    "weights": w.weights}],
                  [{"dataset": 0, "parent": ta.parent, "n_rates": len(w.rates)}] * n_chains)
    for i in range(n_chains):
        b.set_model(i, model)
        b.set_prior_exponential(i, prior)
    b.use_graph(True)

    class Recording(S.EngineLeapfrog):
        """Exact leapfrog that keeps (x, logL, dlogL/dx) of its steps."""
        on = False

        def __call__(self, reqs):
            out = super().__call__(reqs)
            if self.on:
                for q, st, ll in zip(reqs, out, self.last_ll):
                    if q[3] != 0.0:   # molifier off: g = dlogL/dx - 1/prior
                        self.pairs.append((st.x, float(ll), st.g + 1.0 / prior))
            return out

    lf = Recording(b, n_chains, [d] * n_chains)
    lf.pairs = []
    engine_leapfrog = b.leapfrog

    def leapfrog_keep_ll(ed, x, r, g):
        lp, ll = engine_leapfrog(ed, x, r, g)
        lf.last_ll = ll
        return lp, ll
    b.leapfrog = leapfrog_keep_ll
    blv = mcp.get_branchlength_vector(w.tree)
    warm = [S.NUTSChain(blv, np.random.default_rng(100 + i), target=0.8, tree_depth=depth,
                        molifier=False) for i in range(n_chains)]

    def tr(i, k):
        if i == 0 and k == n_warm // 3:
            lf.on = True
        if i == 0 and k % 25 == 0:
            log(f"  warm-up iteration {k}")
    t0 = time.perf_counter()
    S.run_nuts(warm, lf, 0, n_adapt=n_warm, trace=tr)
    warm_s = time.perf_counter() - t0
    lf.on = False
    X = np.array([p[0] for p in lf.pairs])
    L = np.array([p[1] for p in lf.pairs])
    G = np.array([p[2] for p in lf.pairs])
    log(f"  warm-up {warm_s:.1f} s, {len(L)} training pairs")

    def summary(draws, lps, wall):
        tl = np.concatenate([c.sum(axis=1) for c in draws])
        ess_tl, ess_lp = S.ess_bm(tl, 25), S.ess_bm(np.concatenate(lps), 25)
        ess_b = min(S.ess_bm(np.concatenate([c[:, k] for c in draws]), 25) for k in range(d))
        return {"wall_s": wall, "ess_tree_length": ess_tl, "ess_logpost": ess_lp,
                "min_ess_branch": ess_b, "ess_per_s": min(ess_tl, ess_lp) / wall,
                "mean_tree_length": float(tl.mean())}

    x_warm = [c.x for c in warm]
    eps_warm = np.array([c.epsilon for c in warm])
    plain = S.NativeNUTS(x_warm, [200 + i for i in range(n_chains)], target=0.8,
                         tree_depth=depth, molifier=False)
    plain.set_state(eps=eps_warm)
    plain.use_engine(b)
    dp, lp_mat, wall_p, rounds_p, lfs_p = plain.run(n_iter)
    lpp = [lp_mat[:, i] for i in range(n_chains)]
    out = {"workload": f"{len(ta.leaf_nums)} taxa, 4-state GTR, {w.C} weighted patterns, "
                       f"4 gamma cats, exponential({prior}) branch-length prior, {n_chains} chains "
                       f"in lockstep (native driver), NUTS target 0.8, tree depth {depth}, "
                       "molifier off",
           "warmup_iterations": n_warm, "iterations": n_iter, "training_pairs": int(len(L)),
           "plain": summary(dp, lpp, wall_p)}
    out["plain"]["exact_leapfrogs"] = int(lfs_p)
    log(f"  plain: {wall_p:.1f} s, ESS/s {out['plain']['ess_per_s']:.2f}")
    s = mcp.Surrogate(d, 256, 0, 1)
    s.fit_scaling(X, L)
    t1 = time.perf_counter()
    s.set_data(X, L, grad=G)
    losses = s.train(epochs, batch=256, lr=1e-3, lr_final=1e-5, seed=2, grad_weight=grad_weight)
    train_s = time.perf_counter() - t1
    Dp = np.concatenate(dp)
    err = s.predict(Dp) - (np.concatenate(lpp) - S.exponential_logprior(Dp, prior))
    s.use_graph(True)
    da = S.NativeNUTS(x_warm, [300 + i for i in range(n_chains)], target=0.8, tree_depth=depth,
                      molifier=False)
    da.set_state(eps=eps_warm)
    da.use_surrogate(s, prior)

    def exact_logpost(Xs):
        return b.logl(Xs.reshape(-1)) + S.exponential_logprior(Xs, prior)
    dd, lp_da, wall_d, st = S.run_da_nuts_native(da, s, prior, exact_logpost, n_iter)
    lpd = [lp_da[:, i] for i in range(n_chains)]
    out["da"] = summary(dd, lpd, wall_d)
    out["da"].update({"second_stage_accept": st["accept_rate"],
                      "exact_evaluations": int(st["exact_evals"]),
                      "surrogate_leapfrog_rounds": int(st["rounds"])})
    A = np.concatenate(dd)
    z = np.array([abs(A[:, k].mean() - Dp[:, k].mean())
                  / np.hypot(S.mcse_bm(A[:, k], 25), S.mcse_bm(Dp[:, k], 25)) for k in range(d)])
    out["surrogate"] = {"network": f"{d} -> 256 -> 256 -> 1 tanh MLP + separable quadratic skip, "
                                   "BF16 MFMA / FP32 accumulation",
                        "training": f"{epochs} epochs Adam, batch 256, value + gradient matching "
                                    f"(weight {grad_weight:g})",
                        "train_s": train_s, "final_loss": float(losses[-1]),
                        "error_sd_on_plain_draws": float(err.std()),
                        "error_mean_on_plain_draws": float(err.mean())}
    out["posterior_agreement"] = {"branches": d, "frac_abs_z_gt_3": float(np.mean(z > 3)),
                                  "max_abs_z": float(z.max())}
    out["da_over_plain_ess_per_s"] = out["da"]["ess_per_s"] / out["plain"]["ess_per_s"]
    out["da_over_plain_including_training"] = (
        min(out["da"]["ess_tree_length"], out["da"]["ess_logpost"]) / (wall_d + train_s)
        / out["plain"]["ess_per_s"])
    b.close()
    return out


def train_flops(rows, d, H, grad_matching):
    """Algorithmic matrix-product flops of `rows` training rows of the
    d -> H -> H -> 1 network: forward, parameter backward and, with gradient
    matching, the second backward pass through the input gradient."""
    if grad_matching:
        return 2.0 * rows * H * (5 * d + 6 * H)
    return 2.0 * rows * H * (2 * d + 3 * H)


def bench_train(d=398, H=256, n=65536, epochs=4, batches=(256, 1024), grad_weight=1e4):
    """Surrogate training throughput (BF16 MFMA, FP32 accumulation) on a
    synthetic smooth target; the kernels' work does not depend on the values."""
    rng = np.random.default_rng(20261202)
    X = np.exp(rng.normal(-2.5, 0.5, (n, d)))
    A = rng.normal(0, 1.0 / np.sqrt(d), (d, 8))
    L = np.log(X)
    u = (L + 2.5) @ A
    y = -5000.0 - 0.5 * np.sum(u * u, axis=1)
    G = -(u @ A.T) / X
    out = {"workload": f"{d} inputs -> {H} -> {H} -> 1 tanh MLP + separable quadratic skip, "
                       f"{n} resident training rows, value + gradient matching "
                       f"(weight {grad_weight:g}), Adam",
           "flops_per_row": train_flops(1, d, H, True), "runs": []}
    for batch in batches:
        s = mcp.Surrogate(d, H, 0, 1)
        s.fit_scaling(X, y)
        s.set_data(X, y, G)
        s.train(1, batch=batch, lr=1e-3, lr_final=1e-3, seed=1, grad_weight=grad_weight)
        t0 = time.perf_counter()
        losses = s.train(epochs, batch=batch, lr=1e-3, lr_final=1e-4, seed=2,
                         grad_weight=grad_weight)
        dt = time.perf_counter() - t0
        s.train_graph(False)
        t1 = time.perf_counter()
        s.train(1, batch=batch, lr=1e-4, lr_final=1e-4, seed=3, grad_weight=grad_weight)
        dt_eager = time.perf_counter() - t1
        steps = epochs * (-(-n // batch))
        tf = train_flops(n, d, H, True) * epochs / dt / 1e12
        out["runs"].append({"batch": batch, "epochs": epochs, "wall_s": dt,
                            "ms_per_step": 1e3 * dt / steps, "rows_per_s": n * epochs / dt,
                            "tflops": tf, "eager_ms_per_step": 1e3 * dt_eager / (steps / epochs),
                            "final_loss": float(losses[-1])})
        log(f"  train batch {batch}: {1e3 * dt / steps:.3f} ms/step, {tf:.2f} TFLOP/s "
            f"(eager {1e3 * dt_eager / (steps / epochs):.3f} ms/step)")
        s.close()
    best = max(out["runs"], key=lambda r: r["tflops"])
    out["roofline"] = {"bound": "mfma", "achieved": best["tflops"], "peak": 2500.0,
                       "unit": "TFLOP/s", "frac": best["tflops"] / 2500.0, "traffic": None,
                       "batch": best["batch"]}
    return out


def bench_dense(n_sites, steps, cpu: bool, cpu_seconds: float):
    """The dense 20-state workload (1000 taxa, n_sites sites as leaf state
    codes, 4 gamma categories) on the FP64-MFMA level kernels; the CPU oracle
    is timed on a 1000-site slice (the full size needs > 300 GB host memory
    in the oracle's layout)."""
    t0 = time.perf_counter()
    w = mcp.synthetic.make_dense_workload(C=n_sites)
    log(f"  dense workload generated in {time.perf_counter() - t0:.1f} s")
    b = mcp.Batch([w.dataset()], [{"dataset": 0, "parent": w.parent, "n_rates": w.R}])
    b.set_model(0, mcp.Model(w.U, w.D, w.Uinv, w.mu, w.rates, w.pi))
    b.set_schedule("levels")
    b.logl_grad(w.blv)
    t1 = time.perf_counter()
    for _ in range(steps):
        b.logl_grad(w.blv)
    wall = time.perf_counter() - t1
    b.set_timing(True)
    b.logl_grad(w.blv)
    ms, ph = b.launch_times()
    b.set_timing(False)
    bytes_fwd, bytes_adj = b.algorithmic_bytes()
    b.close()
    fwd_ms, adj_ms = float(np.sum(ms[ph == 2])), float(np.sum(ms[ph == 3]))
    n = len(w.leaf_nums)
    # useful FP64 flops: 2 S^2 per (pattern, category) and internal child edge
    # in the post-order pass, ~8 S^2 per internal and 4 S^2 per leaf branch in
    # the pre-order pass
    fl_fwd = 2.0 * w.S ** 2 * w.R * w.C * (2 * (n - 1) - n)
    fl_adj = 1.0 * w.S ** 2 * w.R * w.C * (8 * (n - 2) + 4 * n)
    out = {"workload": f"{n} taxa (random coalescence), t ~ Exp(0.05), 20 states "
                       "(exchangeabilities iid Exp(1), pi ~ Dir(10)), 4 gamma cats (alpha 0.5), "
                       f"{w.C} simulated sites as uint8 tip codes, FP64 logL + full gradient",
           "sites": w.C, "nodes": w.N, "steps": steps,
           "ms_per_step": 1e3 * wall / steps, "value": w.C * steps / wall,
           "unit": "site-pattern evals/s",
           "post_order_ms": fwd_ms, "pre_order_ms": adj_ms,
           "post_order_GBs": bytes_fwd / (fwd_ms * 1e-3) / 1e9,
           "pre_order_GBs": bytes_adj / (adj_ms * 1e-3) / 1e9,
           "algorithmic_GB_per_step": (bytes_fwd + bytes_adj) / 1e9,
           "post_order_fp64_TFLOPs": fl_fwd / (fwd_ms * 1e-3) / 1e12,
           "pre_order_fp64_TFLOPs": fl_adj / (adj_ms * 1e-3) / 1e12,
           "fp64_mfma_peak_TFLOPs": 78.6}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = min(os.cpu_count() or 1, 16)
        oracle.set_threads(threads)
        cols = slice(0, 1000)
        x = w.tip_partials(cols)
        ta = mcp.tree_arrays(mcp.tree_from_parent(w.parent, w.blv))
        k, t2 = 0, time.perf_counter()
        while True:
            oracle.logl_grad(ta, x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, w.blv)
            k += 1
            dt = time.perf_counter() - t2
            if dt >= cpu_seconds:
                break
        out["cpu_baseline"] = {"value": 1000 * k / dt, "unit": "site-pattern evals/s",
                               "cores": threads, "kind": "port",
                               "sample": f"{k} logL+grad evaluations of the first 1000 sites "
                                         f"with the C oracle, {threads} OpenMP threads"}
    return out


WORKLOADS = {
    "gtr100": ("100 taxa random topology, t~Exp(0.05), 4-state GTR, 4 gamma cats (alpha 0.5), "
               "10,000 weighted patterns, FP64 logL + d logL/d t"),
}


def cpu_baseline(mcp, w, blv, budget_s):
    """The CPU oracle (reference-algorithm restatement) on this host's cores,
    timed on the same workload for about budget_s seconds."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = min(os.cpu_count() or 1, 16)
    oracle.set_threads(threads)
    oracle.logl_grad(w.ta, w.x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, blv, weights=w.weights)
    n, t0 = 0, time.perf_counter()
    while True:
        oracle.logl_grad(w.ta, w.x, w.U, w.D, w.Uinv, w.mu, w.rates, w.pi, blv, weights=w.weights)
        n += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s and n >= 2:
            break
    return {"value": w.C * n / dt, "unit": "site-pattern evals/s", "cores": threads,
            "kind": "port",
            "sample": f"{n} full logL+grad evaluations of the same workload "
                      f"({w.C} patterns, {w.N} nodes) with the C oracle, {threads} OpenMP threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="gtr100", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-trees", type=int, default=256,
                    help="trees of the ragged batch (4 chains each); 0 skips it.  The batch "
                         "is generated in spawned worker processes before the GPU is touched; "
                         "use 0 under a profiler that initialises the GPU at load")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--batch-workers", type=int, default=8,
                    help="spawned generator processes (0: generate in this process)")
    ap.add_argument("--nuts-iters", type=int, default=200,
                    help="post-adaptation NUTS iterations of the ESS/s section (0 skips it)")
    ap.add_argument("--only-batched", action="store_true",
                    help="run only the ragged-batch section (for profiling it alone)")
    ap.add_argument("--schedule", default="auto", choices=["auto", "levels", "hybrid", "fused"],
                    help="launch schedule of the timed evaluations (auto: autotuned)")
    ap.add_argument("--hybrid-width", type=int, default=4,
                    help="widest level fused into the top launch with --schedule hybrid")
    ap.add_argument("--cols-per-thread", type=int, default=0,
                    help="fixed tile width of the level kernels (0: autotuned); with a fixed "
                         "--schedule this pins the configuration a profiler sees")
    ap.add_argument("--da-iters", type=int, default=100,
                    help="iterations of the delayed-acceptance section (200-taxon GTR); 0 skips")
    ap.add_argument("--dense-sites", type=int, default=100_000,
                    help="sites of the dense 20-state section (1000 taxa); 0 skips it")
    ap.add_argument("--dense-steps", type=int, default=5)
    ap.add_argument("--only-dense", action="store_true",
                    help="run only the dense 20-state section (for profiling it alone)")
    ap.add_argument("--train-epochs", type=int, default=4,
                    help="epochs of the surrogate-training section (0 skips it)")
    ap.add_argument("--only-train", action="store_true",
                    help="run only the surrogate-training section (for profiling it alone)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path shards over site patterns across GPUs by design, not built here)",
              file=sys.stderr)
        sys.exit(2)

    batch_data = None
    if args.batch_trees > 0:
        log(f"generating ragged batch ({args.batch_trees} trees x 4 chains)")
        batch_data = make_batch_data(args.batch_trees,
                                     min(args.batch_workers, os.cpu_count() or 1))
    if mcp._lib.device_count() < 1:
        log("no HIP device visible")
        sys.exit(1)
    if args.only_dense:
        res = bench_dense(args.dense_sites, args.dense_steps, not args.no_cpu,
                          min(args.cpu_seconds, 8.0))
        print(json.dumps({"dense20": res}), flush=True)
        return
    if args.only_train:
        res = bench_train(epochs=max(1, args.train_epochs))
        print(json.dumps({"surrogate_training": res}), flush=True)
        return
    if args.only_batched:
        if batch_data is None:
            log("--only-batched needs --batch-trees > 0")
            sys.exit(2)
        res = bench_batched(batch_data, args.batch_steps, 2, 16, args.schedule)
        print(json.dumps({"batched": res}), flush=True)
        return
    log(f"building workload {args.workload}")
    w = mcp.synthetic.make_workload(args.workload)
    blv = mcp.get_branchlength_vector(w.tree)
    plan = mcp.Plan(w.ta.parent, w.x, len(w.rates), weights=w.weights)
    plan.set_model(mcp.Model(w.U, w.D, w.Uinv, w.mu, w.rates, w.pi))
    plan.set_schedule("levels" if args.schedule == "auto" else args.schedule,
                      args.hybrid_width if args.schedule == "hybrid" else None)
    plan.logl_grad(blv)
    if args.cols_per_thread > 0:
        n_states = int(np.asarray(w.U).shape[0])
        plan.set_tiles(n_states, args.cols_per_thread)
        tiles, tune_rows = {n_states: plan.tiles(n_states)}, []
        log(f"fixed tiles (columns per thread by state count): {tiles}")
    else:
        tiles, tune_rows = plan.autotune(reps=5)
        log(f"autotuned tiles (columns per thread by state count): {tiles}")
    sched_rows = []
    if args.schedule == "auto":
        _, sched_rows = plan.autotune_schedule(reps=9)
        if args.cols_per_thread <= 0:
            # the best tile width can differ under the chosen schedule (the
            # fused launches have one block per tile): tune it again there,
            # then the schedule once more
            tiles, tune_rows = plan.autotune(reps=5)
            _, sched_rows = plan.autotune_schedule(reps=9)
            log(f"tiles under the chosen schedule: {tiles}")
    sched, hwidth = plan.schedule()
    log(f"schedule: {sched}" + (f" (hybrid width {hwidth})" if sched == "hybrid" else ""))
    for _ in range(args.warmup):
        plan.logl_grad(blv)
    log(f"warmup done ({args.warmup}); timing {args.steps} steps")

    # timed region: eager launches, no instrumentation
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.logl_grad(blv)
    wall = time.perf_counter() - t0
    value = w.C * args.steps / wall
    log(f"eager: {1e3 * wall / args.steps:.3f} ms/step")

    # the same steps again with every kernel bracketed by HIP events on the
    # engine's stream: per-launch durations and algorithmic bytes
    plan.set_timing(True)
    launch_ms = None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.logl_grad(blv)
        ms, ph = plan.launch_times()
        launch_ms = ms if launch_ms is None else launch_ms + ms
    lbytes = plan.launch_bytes()
    iwall = time.perf_counter() - t0
    plan.set_timing(False)
    bytes_fwd, bytes_adj = plan.algorithmic_bytes()
    launch_ms = launch_ms / args.steps
    per_phase = {k: float(np.sum(launch_ms[ph == k])) for k in (1, 2, 3, 4)}
    n_phase = {k: int(np.sum(ph == k)) for k in (1, 2, 3, 4)}
    dom = 3 if per_phase[3] >= per_phase[2] else 2
    kbytes, kms, nl = float(np.sum(lbytes[ph == dom])), per_phase[dom], n_phase[dom]
    cpt = tiles.get(w.S, 0)
    pass_name = "pre-order" if dom == 3 else "post-order"
    kname = f"{pass_name} pass: k_level<{w.S},{cpt},{'true' if dom == 3 else 'false'}>"
    if sched == "hybrid":
        kname = (f"{pass_name} pass: {nl} launch(es) -- k_level<{w.S},{cpt},true> per wide "
                 f"level and k_fused<{w.S},{cpt},true> over the narrow levels (the first fused "
                 "launch also walks the narrow top of the post-order pass)")
    elif sched == "fused":
        kname = f"k_fused<{w.S},{cpt},true> (both passes, one launch)"
    achieved = kbytes / (kms * 1e-3) / 1e9

    # graph replay of the same evaluation
    plan.use_graph(True)
    for _ in range(3):
        plan.logl_grad(blv)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        plan.logl_grad(blv)
    gwall = time.perf_counter() - t1
    plan.use_graph(False)
    gvalue = w.C * args.steps / gwall
    log(f"graph: {1e3 * gwall / args.steps:.3f} ms/step")

    # the plain per-level schedule, eager, for reference
    plan.set_schedule("levels")
    for _ in range(3):
        plan.logl_grad(blv)
    t2 = time.perf_counter()
    for _ in range(args.steps):
        plan.logl_grad(blv)
    owall = time.perf_counter() - t2
    log(f"levels schedule: {1e3 * owall / args.steps:.3f} ms/step")

    result = {
        "metric": "logL+grad site-pattern evals/sec",
        "value": value,
        "unit": "site-pattern evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {WORKLOADS[args.workload]}",
                   "taxa": int(len(w.ta.leaf_nums)), "nodes": int(w.N), "states": int(w.S),
                   "patterns": int(w.C), "rate_categories": int(len(w.rates)),
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": kname, "launches_per_step": nl,
                     "algorithmic_bytes_per_step": kbytes,
                     "kernel_ms_per_step": kms,
                     "avg_launch_us": 1e3 * kms / max(nl, 1)},
        "schedule": sched,
        "hybrid_width": hwidth if sched == "hybrid" else None,
        "cols_per_thread": {str(k): v for k, v in tiles.items()},
        "tiles_autotuned": args.cols_per_thread <= 0,
        "autotune_ms": [list(r) for r in tune_rows],
        "autotune_schedule_ms": [list(r) for r in sched_rows],
        "graph_value": gvalue,
        "instrumented_ms_per_step": 1e3 * iwall / args.steps,
        "kernel_ms_per_step": {"P_build": per_phase[1], "post_order": per_phase[2],
                               "pre_order": per_phase[3], "reduce": per_phase[4]},
        "launches_per_step": n_phase,
        "algorithmic_bytes_per_step": {"post_order": bytes_fwd, "pre_order": bytes_adj},
        "levels_schedule": {"value": w.C * args.steps / owall,
                            "ms_per_step": 1e3 * owall / args.steps},
    }
    if batch_data is not None:
        log("timing ragged batch")
        result["batched"] = bench_batched(batch_data, args.batch_steps, 2, 16, args.schedule)
        result["batched"]["workload"] = (
            f"{args.batch_trees} trees, taxa ~ U{{16..512}}, patterns ~ LogU[500, 20000], "
            "restriction/GTR-4 50/50, 4 gamma cats, 4 chains per tree, FP64 logL+grad")
        log(f"batched: {result['batched']['ms_per_step']:.2f} ms/step")
    if args.nuts_iters > 0:
        log("NUTS ESS/s section")
        result["nuts"] = bench_nuts(100, args.nuts_iters, not args.no_cpu)
        log(f"nuts: gpu {result['nuts']['gpu']['ess_per_s']:.1f} ESS/s")
    if args.da_iters > 0:
        log("delayed-acceptance section")
        result["delayed_acceptance"] = bench_da(100, args.da_iters)
        log(f"DA: {result['delayed_acceptance']['da_over_plain_ess_per_s']:.2f}x plain ESS/s")
    if args.train_epochs > 0:
        log("surrogate-training section")
        result["surrogate_training"] = bench_train(epochs=args.train_epochs)
    if args.dense_sites > 0:
        log("dense 20-state section")
        result["dense20"] = bench_dense(args.dense_sites, args.dense_steps, not args.no_cpu,
                                        min(args.cpu_seconds, 8.0))
        log(f"dense20: {result['dense20']['ms_per_step']:.1f} ms/step")
    if not args.no_cpu:
        log("timing CPU baseline")
        result["cpu_baseline"] = cpu_baseline(mcp, w, blv, args.cpu_seconds)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
