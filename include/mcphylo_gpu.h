/*
 * mcphylo_gpu.h -- C ABI of the MI355X (gfx950) pruning likelihood engine.
 *
 * Drop-in boundary for MCPhylo.jl's PhyloDist hot path: the engine replaces
 * the body of
 *   logpdf(d::PhyloDist, x)         src/distributions/Phylodist.jl:107-122
 *   gradlogpdf(d::PhyloDist, x)     src/distributions/Phylodist.jl:124-138
 *   logpdf(d::MultiplePhyloDist, x) src/distributions/Phylodist.jl:281-288
 *   __logpdf(d::MultiplePhyloDist)  src/distributions/Phylodist.jl:290-297
 * i.e. FelsensteinFunction (src/Likelihood/LikelihoodCalculator_Node.jl:3-114)
 * and the loop nests of src/Likelihood/VectorizedFunctions.jl:2-213.
 *
 * Conventions (those of the reference):
 *  - nodes are numbered 1..N, the root is N; a parent array gives the parent
 *    num of node num i+1 (0 for the root);
 *  - branch lengths and gradients are vectors of length N-1 indexed by num;
 *  - tip partials use the reference layout x[s, c, v] (column-major: state
 *    fastest, then site/pattern, then node);
 *  - U, Uinv are S x S column-major; Q = U diag(D) Uinv; P(t) =
 *    U diag(exp(mu * t * D * rate)) Uinv (VectorizedFunctions.jl:116-168);
 *  - with R > 1 categories the log-likelihood is the sum over every
 *    (pattern, category) pair, as in VectorizedFunctions.jl:37-40,76-87;
 *  - numeric failure (log 0, NaN inputs) is returned as -Inf/NaN values,
 *    never as an error status (reference: simulation.jl:246,263-267).
 * All pointers are host memory owned by the caller; nothing is retained
 * after a call returns.  A handle is used by one host thread at a time;
 * distinct handles may be used concurrently (one HIP stream each).
 */
#ifndef MCPHYLO_GPU_H
#define MCPHYLO_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mcp_engine mcp_plan;  /* one tree: a batch of one item */
typedef struct mcp_engine mcp_batch; /* many (tree, chain) items */

/* Status codes; the Julia wrapper maps them to ArgumentError,
 * DimensionMismatch and ErrorException. */
enum {
    MCP_OK = 0,
    MCP_ERR_INVALID = 1,     /* bad argument (null pointer, malformed tree) */
    MCP_ERR_DIMENSION = 2,   /* sizes that do not agree (DimensionMismatch) */
    MCP_ERR_COMPLEX = 3,     /* non-finite model input */
    MCP_ERR_NOMEM = 4,       /* device or host allocation failed */
    MCP_ERR_HIP = 5,         /* HIP runtime error */
    MCP_ERR_UNSUPPORTED = 6, /* state count or tip alphabet outside the built kernels */
    MCP_ERR_STATE = 7        /* model not set before evaluation */
};

/* Merge identical alignment columns into weighted patterns (the reference
 * has no compression; sums are unchanged up to summation order). */
#define MCP_FLAG_COMPRESS_PATTERNS 1u

/* ---- single tree (PhyloDist) -------------------------------------------- */

typedef struct mcp_plan_desc {
    int32_t n_nodes;                /* N (root = N) */
    int32_t n_states;               /* S */
    int32_t n_patterns;             /* C (sites, or patterns with weights) */
    int32_t n_rates;                /* R rate categories */
    const int32_t* parent;          /* [N] */
    const double* tip_partials;     /* [S*C*N] x[s,c,v]; only leaf slices are read */
    const double* pattern_weights;  /* [C] or NULL (all ones) */
    int32_t device;                 /* HIP device ordinal */
    uint32_t flags;                 /* MCP_FLAG_* */
} mcp_plan_desc;

/* Replaces the per-call setup of Phylodist.jl:108-120 (post_order, leaf
 * lookup, my_repeat, deepcopy): uploads the tips once as state codes. */
int mcp_plan_create(mcp_plan** out, const mcp_plan_desc* desc);
void mcp_plan_destroy(mcp_plan* plan);

/* The eigensystem d.substitution_model(base_freq, rates) returns
 * (Phylodist.jl:110-113), the category rates d.rates and base_freq.
 * U, Uinv: [S*S] column-major; D, pi: [S]; rates: [R]. */
int mcp_set_model(mcp_plan* plan, const double* U, const double* D, const double* Uinv,
                  double mu, const double* rates, const double* pi);

/* New topology over the same leaves (after an NNI move, refraction.jl:65). */
int mcp_set_topology(mcp_plan* plan, const int32_t* parent);

/* logpdf(d::PhyloDist, x) (Phylodist.jl:107-122): blv [N-1] by num. */
int mcp_logl(mcp_plan* plan, const double* blv, double* ll);

/* gradlogpdf(d::PhyloDist, x) (Phylodist.jl:124-138): ll and grad [N-1]. */
int mcp_logl_grad(mcp_plan* plan, const double* blv, double* ll, double* grad);

/* ---- ragged batch of (tree, chain) items (MultiplePhyloDist and beyond) -- */

typedef struct mcp_dataset_desc {
    int32_t n_states;
    int32_t n_patterns;
    int32_t n_leaves;
    const int32_t* leaf_nums;       /* [n_leaves] node nums of the leaves */
    const double* tip_partials;     /* [S*C*n_leaves] x[s,c,j] for leaf leaf_nums[j], or NULL */
    const double* pattern_weights;  /* [C] or NULL */
    uint32_t flags;                 /* MCP_FLAG_* */
    /* Alternative to tip_partials (used when tip_partials is NULL): state
     * codes tip_codes[j*C + c] indexing rows of code_table [n_codes x S]
     * (e.g. S one-hot rows plus an all-ones row for gaps). */
    const uint8_t* tip_codes;
    int32_t n_codes;
    const double* code_table;
} mcp_dataset_desc;

typedef struct mcp_item_desc {
    int32_t dataset;                /* index into mcp_batch_desc.datasets */
    int32_t n_nodes;
    int32_t n_rates;
    const int32_t* parent;          /* [n_nodes]; leaves must be the dataset's leaf_nums */
} mcp_item_desc;

typedef struct mcp_batch_desc {
    int32_t n_datasets;
    const mcp_dataset_desc* datasets;
    int32_t n_items;
    const mcp_item_desc* items;
    int32_t device;
    int32_t waves_per_block;        /* wavefronts per block of the fused schedule, 1..8;
                                       0 = default (4) */
} mcp_batch_desc;

/* MultiplePhyloDist (Phylodist.jl:143-272) generalised to uneven trees,
 * state counts, pattern counts and per-chain branch lengths. */
int mcp_batch_create(mcp_batch** out, const mcp_batch_desc* desc);
void mcp_batch_destroy(mcp_batch* batch);
int mcp_batch_set_model(mcp_batch* batch, int32_t item, const double* U, const double* D,
                        const double* Uinv, double mu, const double* rates, const double* pi);
int mcp_batch_set_topology(mcp_batch* batch, int32_t item, const int32_t* parent);

/* blv_concat: the items' branch-length vectors back to back (N_i - 1 each,
 * item order); ll: [n_items]; grad_concat: same layout as blv_concat.
 * logpdf(::MultiplePhyloDist) (Phylodist.jl:281-288) is the sum of ll;
 * __logpdf (Phylodist.jl:290-297) is the per-tree (ll, grad). */
int mcp_batch_logl(mcp_batch* batch, const double* blv_concat, double* ll);
int mcp_batch_logl_grad(mcp_batch* batch, const double* blv_concat, double* ll,
                        double* grad_concat);

/* Per-(branch, category) components of the gradient: grad_vr [N-1][R]
 * (category fastest; per item, concatenated for a batch), so that
 * grad[v] = sum_r grad_vr[v][r].  Because P of branch v in category r
 * depends on rho_r t_v only, d logL / d rho_r = sum_v t_v grad_vr[v][r] / rho_r
 * -- the gradient with respect to the rate parameters (and, through
 * discrete_gamma_rates, Rates.jl:11-37, the gamma shape).  No reference
 * counterpart: FelsensteinFunction/6 returns branch gradients only. */
int mcp_logl_grad_by_rate(mcp_plan* plan, const double* blv, double* ll, double* grad_vr);
int mcp_batch_logl_grad_by_rate(mcp_batch* batch, const double* blv_concat, double* ll,
                                double* grad_vr_concat);

/* Model gradient: besides (ll, grad), dlogl_dP [N-1][R][S][S] = d logL /
 * d P_{v,r}[s][s1] for every branch and category (row s = parent state, the
 * layout of P) and dlogl_dpi [S] = d logL / d pi_s through the root term
 * log sum_s pi_s Ft_s (VectorizedFunctions.jl:76-87).  With the substitution
 * model's eigensystem (P = U diag(exp(mu rho t D)) Uinv,
 * SubstitutionModels.jl:62-74) they give the gradient with respect to the
 * model's own parameters (GTR exchangeabilities and base frequencies) by the
 * chain rule on the host.  Per item, concatenated for a batch.  Needs S <= 8
 * (MCP_ERR_UNSUPPORTED otherwise); runs every pre-order level as its own
 * launch whatever the schedule.  No reference counterpart: the reference's
 * gradient covers branch lengths only (FelsensteinFunction/6,
 * LikelihoodCalculator_Node.jl). */
int mcp_logl_grad_model(mcp_plan* plan, const double* blv, double* ll, double* grad,
                        double* dlogl_dP, double* dlogl_dpi);
int mcp_batch_logl_grad_model(mcp_batch* batch, const double* blv_concat, double* ll,
                              double* grad_concat, double* dlogl_dP_concat,
                              double* dlogl_dpi_concat);

/* ---- device leapfrog (NUTS / HMC over branch lengths) ------------------- */

/* Branch-length prior added to the log posterior of item `item` by the
 * leapfrog: kind 0 none, kind 1 exponentialBL(scale) -- each branch
 * Exponential with mean `scale` (src/Likelihood/Prior.jl:50-56,80-83). */
int mcp_hmc_set_prior(mcp_batch* batch, int32_t item, int32_t kind, double scale);

/* General form: kind 0 none (nparams 0); kind 1 exponentialBL
 * (params = {scale}); kind 2 CompoundDirichlet (params = {alpha, a, beta,
 * c}, src/Likelihood/Prior.jl:1-37 with the field order of
 * CompoundDirichlet(alpha, a, beta, c), src/distributions/TreeDistribution.jl:23-39):
 *   logprior = alpha log beta - lgamma(alpha) - beta T - lgamma(a) - lgamma(c)
 *              + lgamma(a + c) + (a - 1) sum_leaf log t + (a c - 1) sum_int log t
 *              + (alpha - a n_leaf - a c (n_leaf - 3)) log T,   T = sum t,
 * where leaf branches are those above leaves and internal the others. */
int mcp_hmc_set_prior_params(mcp_batch* batch, int32_t item, int32_t kind, const double* params,
                             int32_t nparams);

/* One leapfrog step for every item, as refraction! does it
 * (src/samplers/tree_hamiltonian/refraction.jl:2-31; reflection at zero in
 * time order as in ref_NNI!, :35-91, without the NNI topology attempt):
 *   r += eps/2 g;  y = x + eps r, reflected at zero;  x = molifier(y, delta);
 *   (ll, grad) = logL+grad(x);  g = (grad + dlogprior) * scale_fac(x, delta);
 *   r += eps/2 g;  logpost = ll + logprior(x)
 * (molifier / scale_fac: src/samplers/tree_hamiltonian/auxilliary.jl:113-119).
 * eps_delta: [2*n_items] (eps, delta) pairs; x, r, g: in/out, blv_concat
 * layout; logpost: out [n_items]; ll: out [n_items] or NULL.  The whole step
 * is one HIP graph replay when mcp_use_graph is on. */
int mcp_hmc_leapfrog(mcp_batch* batch, const double* eps_delta, double* x, double* r, double* g,
                     double* logpost, double* ll);

/* ---- launch control ----------------------------------------------------- *
 * No reference counterpart for this section: FelsensteinFunction
 * (LikelihoodCalculator_Node.jl:3-114) walks the nodes one after another and
 * threads each node's loops with @tturbo; here the launch schedule, tile
 * widths, graphs and timing are the engine's own. */

/* Capture the whole evaluation (upload of branch lengths, P/dP build, every
 * level, reductions, download) into HIP graphs that later calls replay; a
 * topology change re-captures.  on = 0 switches back to eager launches.
 * Applies to plans and batches (mcp_plan and mcp_batch are one type). */
int mcp_use_graph(mcp_batch* batch, int32_t on);

/* Launch schedule of an evaluation.
 *  MCP_SCHEDULE_FUSED   one launch walks every level of a (tree, category,
 *                       pattern tile) with block barriers between levels;
 *  MCP_SCHEDULE_LEVELS  one launch per post-order level (nodes of equal
 *                       height) and per pre-order level (nodes of equal
 *                       depth), over every item (default). */
#define MCP_SCHEDULE_FUSED 0
#define MCP_SCHEDULE_LEVELS 1
/*  MCP_SCHEDULE_HYBRID  per-level launches for the wide levels, one fused
 *                       launch for every item's narrow top levels (post-order
 *                       levels above, and pre-order levels from the root
 *                       down to, the first level wider than the hybrid width);
 *                       the narrow levels are consecutive in time, between
 *                       the two passes.  A second fused launch walks the
 *                       deepest pre-order levels when at least two of them,
 *                       and every level below the first, are that narrow. */
#define MCP_SCHEDULE_HYBRID 2
int mcp_set_schedule(mcp_batch* batch, int32_t mode);

/* Widest level (internal nodes of one item) run fused by the HYBRID schedule
 * (default 4). */
int mcp_set_hybrid_width(mcp_batch* batch, int32_t width);

/* Wavefronts per block of the fused launches (1..8; the creation-time
 * waves_per_block, default 4). */
int mcp_set_fused_waves(mcp_batch* batch, int32_t waves);

/* Tile width of the level kernels for state count S: a task is one
 * wavefront over 64 * cpt consecutive patterns (cpt = columns per lane).
 * Built candidates: cpt 1, 2, 4 for S = 2..4 (1 or 2 otherwise); for S = 20,
 * cpt 0 selects the FP64-MFMA kernels and 1 the generic ones.  cpt < 0
 * restores the built default.  The choice changes the summation order of the
 * partial sums (results move in the last bits), so a stored choice can be
 * reapplied with mcp_set_tiles to reproduce results bit for bit.  No
 * reference counterpart (the reference threads its loops with @tturbo). */
int mcp_set_tiles(mcp_batch* batch, int32_t S, int32_t cpt);
int mcp_get_tiles(mcp_batch* batch, int32_t S, int32_t* cpt);

/* Autotuner: for every state count in the engine, time `reps` full logL +
 * gradient evaluations (median, after one untimed) with each built tile
 * width on the engine's own trees and data, and keep the fastest.  Branch
 * lengths: those of the last evaluation (0.1 everywhere before the first).
 * report (may be NULL): up to cap rows (S, cpt, ms); *n_report: rows tried.
 * The model of every item must be set. */
int mcp_autotune(mcp_batch* batch, int32_t reps, double* report, int32_t cap, int32_t* n_report);

/* Autotuner for the launch schedule: times LEVELS, HYBRID with widths 1, 2,
 * 4, 8, 16 and FUSED the same way and keeps the fastest; report rows are
 * (mode, hybrid width, ms).  mcp_get_schedule reads the choice (gradients
 * are bitwise identical across schedules; logL agrees to rounding, because
 * the fused launches sum a tile's log scalers as one product per lane). */
int mcp_autotune_schedule(mcp_batch* batch, int32_t reps, double* report, int32_t cap,
                          int32_t* n_report);
int mcp_get_schedule(mcp_batch* batch, int32_t* mode, int32_t* hybrid_width);

/* Time later evaluations with HIP events on the engine's stream (on != 0).
 * ms[0] whole sequence (first upload to last download); ms[1..4] summed
 * kernel durations of the P/dP build, post-order levels, pre-order levels
 * and reductions -- of the last evaluation.  Under the fused schedule the
 * single fused launch is counted as pre-order (phase 3) when the gradient is
 * computed and as post-order (phase 2) otherwise. */
int mcp_set_timing(mcp_batch* batch, int32_t on);
int mcp_get_timing(mcp_batch* batch, double* ms5);

/* Durations (ms) and phases (1..4 as above) of every kernel launch of the
 * last timed evaluation, in launch order; returns the number of launches
 * (or a negative status).  At most cap entries are written. */
int mcp_get_launch_times(mcp_batch* batch, double* ms, int32_t* phase, int32_t cap);

/* Algorithmic HBM bytes of every launch of the last timed evaluation (same
 * order as mcp_get_launch_times; 0 outside the level passes); returns the
 * number of launches. */
int mcp_get_launch_bytes(mcp_batch* batch, double* bytes, int32_t cap);

/* Algorithmic HBM bytes of one evaluation under this engine's layout:
 * bytes[0] post-order levels, bytes[1] pre-order levels (compulsory slab
 * reads and writes counted once each; tip codes one byte per pattern). */
int mcp_get_bytes(mcp_batch* batch, double* bytes2);

/* Number of items, and total patterns over items (the unit of the
 * site-pattern evals/s metric). */
int mcp_get_sizes(mcp_batch* batch, int64_t* sizes2);

/* ---- learned surrogate of logL (delayed-acceptance first stage) ---------
 * No counterpart in the reference: MCPhylo's samplers call the exact
 * logpdf at every leapfrog (src/samplers/pnuts.jl:300-330 via refraction!).
 * These entry points let a sampler integrate its trajectories on a cheap
 * surrogate and keep the exact likelihood (mcp_logl / mcp_batch_logl) as the
 * second-stage accept/reject, so the posterior is unchanged.
 *
 * Network over d = n_inputs branches (z = standardised log branch lengths):
 *   h1 = tanh(W1 z + b1), h2 = tanh(W2 h1 + b2),
 *   y = w3.h2 + b3 + v.z + q.(z*z),  logL~ = y * y_sd + y_mean,
 *   z_c = (log x_c - x_mean[c]) / x_sd[c].
 * Matrix products on BF16 MFMA with FP32 accumulation; FP32 master weights.
 * Flat FP32 parameter layout (offsets from mcp_mlp_layout, D = d padded to
 * a multiple of 64, H = hidden): W1 [H][D] row-major | b1 [H] | W2 [H][H] |
 * b2 [H] | w3 [H] | b3 [64] (b3[0] used) | v [D] | q [D].  Padding entries
 * are zero and stay zero. */
typedef struct mcp_mlp mcp_mlp;

/* hidden: multiple of 64 (256 for the standard network); Glorot-uniform
 * weights from `seed`, zero biases and skip, identity scaling. */
int mcp_mlp_create(mcp_mlp** out, int32_t n_inputs, int32_t hidden, int32_t device,
                   uint64_t seed);
void mcp_mlp_destroy(mcp_mlp* mlp);

/* offsets9 = {W1, b1, W2, b2, w3, b3, v, q, n_params} */
int mcp_mlp_layout(mcp_mlp* mlp, int64_t* offsets9);
int mcp_mlp_get_params(mcp_mlp* mlp, float* params);
int mcp_mlp_set_params(mcp_mlp* mlp, const float* params);

/* Feature (x_mean, x_sd: [d]) and target (y_mean, y_sd) standardisation;
 * set before mcp_mlp_set_data. */
int mcp_mlp_set_scaling(mcp_mlp* mlp, const float* x_mean, const float* x_sd, float y_mean,
                        float y_sd);

/* Upload a training set of n (state, logL) pairs: X [n][d] branch lengths
 * (row-major, > 0), logl [n], and optionally the exact gradients
 * grad [n][d] = d logL / d x (NULL: none).  It stays resident in device
 * memory. */
int mcp_mlp_set_data(mcp_mlp* mlp, const double* X, const double* logl, const double* grad,
                     int64_t n);

/* Adam (0.9, 0.999, 1e-8) on the loss
 *   (1/2n) sum_r (y_r - t_r)^2 + (grad_weight / 2nd) sum_r |dy/dz_r - G_r|^2
 * of standardised targets t and standardised exact gradients
 * G = dlogL/dx * x * x_sd / y_sd (gradient matching; grad_weight 0 turns it
 * off, > 0 needs gradients in mcp_mlp_set_data).  Minibatches of `batch`
 * rows reshuffled each epoch (seeded); learning rate decaying exponentially
 * from lr to lr_final over the epochs.  epoch_loss: [epochs] mean batch
 * loss, or NULL. */
int mcp_mlp_train(mcp_mlp* mlp, int32_t batch, int32_t epochs, float lr, float lr_final,
                  float grad_weight, uint64_t seed, float* epoch_loss);

/* Gradient of the loss above over the n rows of (X, logl, dlogl), written in
 * the flat parameter layout; loss to *loss (may be NULL).  n <= 65536;
 * dlogl may be NULL when grad_weight is 0. */
int mcp_mlp_param_grad(mcp_mlp* mlp, const double* X, const double* logl, const double* dlogl,
                       int64_t n, float grad_weight, float* grad, float* loss);

/* logL~ [n] and, when grad != NULL, d logL~ / d x [n][d] for the rows of X. */
int mcp_mlp_predict(mcp_mlp* mlp, const double* X, int64_t n, double* logl, double* grad);

/* mcp_hmc_leapfrog on the surrogate posterior logL~(x) + logprior(x) for
 * `chains` chains (x, r, g: [chains][d]; eps_delta: [2*chains]); prior:
 * every branch Exponential(mean prior_scale), none if prior_scale <= 0.
 * Same refraction, molifier and scale_fac steps as mcp_hmc_leapfrog; one HIP
 * graph replay per call when mcp_mlp_use_graph is on. */
int mcp_mlp_leapfrog(mcp_mlp* mlp, int32_t chains, double prior_scale, const double* eps_delta,
                     double* x, double* r, double* g, double* logpost);
int mcp_mlp_use_graph(mcp_mlp* mlp, int32_t on);

/* Training runs each full minibatch as one replay of a captured HIP graph
 * (default on; the minibatch offset and Adam's rate come from a device step
 * counter).  Off launches the same kernels one by one: same results, bit for
 * bit. */
int mcp_mlp_train_graph(mcp_mlp* mlp, int32_t on);

/* ---- native NUTS driver (many chains in lockstep) ------------------------
 * Restates the tree NUTS of src/samplers/pnuts.jl:103-419 (nuts_sub!,
 * buildtree, multinomial sampling, dual averaging) and nutsepsilon
 * (src/samplers/tree_hamiltonian/auxilliary.jl:6-51) for a fixed topology, in
 * C++: one batched leapfrog per round over all chains, so chains of a batch
 * share every launch and no host interpreter sits between the rounds.
 * Chain i owns dims[i] branch lengths at offset sum(dims[0..i)) of the
 * concatenated vectors.  The leapfrog provider is an engine (items =
 * chains; mcp_hmc_leapfrog), a surrogate (mcp_mlp_leapfrog) or a callback
 * with the same contract: for every chain advance (x, r, g) by one leapfrog
 * of (eps, delta) = eps_delta[2i..2i+1] and write the log posterior. */
typedef struct mcp_nuts mcp_nuts;
typedef int (*mcp_leapfrog_fn)(void* ctx, int32_t n_chains, const double* eps_delta, double* x,
                               double* r, double* g, double* logpost);

/* seeds: [n_chains] generator seeds; target: acceptance target of the step
 * size adaptation; molifier: 1 as the reference (delta = 2 eps), 0 for
 * pure reflection at zero. */
int mcp_nuts_create(mcp_nuts** out, int32_t n_chains, const int32_t* dims, const double* x0,
                    const uint64_t* seeds, double target, int32_t tree_depth, int32_t molifier);
void mcp_nuts_destroy(mcp_nuts* nuts);
int mcp_nuts_use_batch(mcp_nuts* nuts, mcp_batch* batch);
int mcp_nuts_use_mlp(mcp_nuts* nuts, mcp_mlp* mlp, double prior_scale);
int mcp_nuts_use_callback(mcp_nuts* nuts, mcp_leapfrog_fn fn, void* ctx);

/* n_iter transitions of every chain (adapt != 0: with step-size adaptation).
 * draws: [n_iter][sum dims] or NULL; logpost: [n_iter][n_chains] or NULL;
 * stats (may be NULL): {leapfrog rounds, leapfrogs summed over chains}. */
int mcp_nuts_run(mcp_nuts* nuts, int32_t n_iter, int32_t adapt, double* draws, double* logpost,
                 int64_t* stats);

/* Current positions, step sizes, log posteriors and last tree depths (any
 * may be NULL); mcp_nuts_set overwrites positions / step sizes / log
 * posteriors (NULL: keep). */
int mcp_nuts_get(mcp_nuts* nuts, double* x, double* eps, double* lf, int32_t* depth);
int mcp_nuts_set(mcp_nuts* nuts, const double* x, const double* eps, const double* lf);

/* Message of the last failed call on this host thread ("" if none). */
const char* mcp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MCPHYLO_GPU_H */
