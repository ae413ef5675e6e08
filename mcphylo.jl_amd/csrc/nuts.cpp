// Native NUTS driver over branch lengths: many chains advanced in lockstep,
// one batched leapfrog per round (exact device leapfrog, surrogate leapfrog
// or a host callback).
//
// Restates the tree sampler of MCPhylo for a fixed topology:
//   nuts_sub! / buildtree (src/samplers/pnuts.jl:179-419): multinomial
//     trajectory sampling, biased progressive sampling at the top level and
//     uniform progressive sampling inside subtrees, divergence check
//     `lu + (logp' - logp0) < 1000`, acceptance statistic of the last doubling;
//   nutsepsilon (src/samplers/tree_hamiltonian/auxilliary.jl:6-51): initial
//     step size search (capped at 60 halvings/doublings);
//   dual averaging (pnuts.jl:145-176) with (gamma, kappa, t0) = (0.05, 0.75,
//     10) and mu = log(10 eps).
// Deliberate differences (as in the host sampler of this package): Euclidean
// U-turn test on branch-length vectors, trajectory ends integrated in
// directions -1 and +1, a NaN acceptance statistic counts as 0, no NNI moves.
// The molifier (auxilliary.jl:113-119) is on by default; with it off the
// leapfrog passes delta = 1e-300 (pure reflection at zero).
//
// Each chain is a C++20 coroutine; a leapfrog request suspends the innermost
// active coroutine of the chain, the driver batches the requests of all
// chains into one leapfrog call and resumes each chain with its result.
#include <math.h>

#include <algorithm>
#include <coroutine>
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <random>
#include <string>
#include <utility>
#include <vector>

#include "mcphylo_gpu.h"

namespace mcp {
int set_last_error(int code, const char* msg);
}
using mcp::set_last_error;

namespace {

struct HState {
    std::vector<double> x, r, g;
    double lf = 0.0;
    double neg_energy() const {
        double rr = 0.0;
        for (double v : r) rr += v * v;
        return lf - 0.5 * rr;
    }
};

struct Chain;

// Coroutine task returning T; awaiting it runs it to completion (its
// leapfrog requests suspend through to the driver) and yields its value.
template <class T>
struct Task {
    struct promise_type {
        T value{};
        std::coroutine_handle<> cont;
        Task get_return_object() {
            return Task{std::coroutine_handle<promise_type>::from_promise(*this)};
        }
        std::suspend_always initial_suspend() noexcept { return {}; }
        struct Final {
            bool await_ready() noexcept { return false; }
            std::coroutine_handle<> await_suspend(std::coroutine_handle<promise_type> h) noexcept {
                auto c = h.promise().cont;
                return c ? c : std::noop_coroutine();
            }
            void await_resume() noexcept {}
        };
        Final final_suspend() noexcept { return {}; }
        void return_value(T v) { value = std::move(v); }
        void unhandled_exception() { std::terminate(); }
    };
    std::coroutine_handle<promise_type> h;
    explicit Task(std::coroutine_handle<promise_type> hh) : h(hh) {}
    Task(Task&& o) noexcept : h(std::exchange(o.h, {})) {}
    Task(const Task&) = delete;
    ~Task() {
        if (h) h.destroy();
    }
    bool await_ready() { return false; }
    std::coroutine_handle<> await_suspend(std::coroutine_handle<> caller) {
        h.promise().cont = caller;
        return h;
    }
    T await_resume() { return std::move(h.promise().value); }
};

struct Meta {
    double alpha = 0.0;
    int nalpha = 0;
};

struct SubTree {
    HState xprime, far, near;
    int n = 0;
    bool s = false;
    double lsw = 0.0;
};

double logaddexp(double a, double b) {
    if (a == -INFINITY) return b;
    if (b == -INFINITY) return a;
    const double m = std::max(a, b);
    return m + std::log(std::exp(a - m) + std::exp(b - m));
}

bool no_uturn(const HState& minus, const HState& plus) {
    double a = 0.0, b = 0.0;
    for (size_t k = 0; k < minus.x.size(); ++k) {
        const double dx = plus.x[k] - minus.x[k];
        a += dx * minus.r[k];
        b += dx * plus.r[k];
    }
    return a >= 0.0 && b >= 0.0;
}

struct Chain {
    int dim = 0;
    size_t off = 0;
    std::vector<double> x;
    std::mt19937_64 rng;
    std::normal_distribution<double> gauss{0.0, 1.0};
    std::uniform_real_distribution<double> unif{0.0, 1.0};
    double target = 0.8;
    int tree_depth = 10;
    bool molifier = true;
    double epsilon = NAN;
    // dual averaging
    int m = 0;
    double s_bar = 0.0, x_bar = 0.0, mu = 0.5;
    const double gamma = 0.05, kappa = 0.75, t0 = 10.0;
    double metro_acc_prob = 0.0;
    bool adapting = false;
    double lf = NAN;
    int depth = 0;
    long long leapfrogs = 0;
    // leapfrog request / result (filled by the driver)
    HState req, res;
    double req_eps = 0.0, req_delta = 0.0;
    bool pending = false;
    std::coroutine_handle<> current;

    double U() { return unif(rng); }
    std::vector<double> randn() {
        std::vector<double> v(dim);
        for (auto& a : v) a = gauss(rng);
        return v;
    }

    // awaitable leapfrog step
    struct Step {
        Chain* c;
        bool await_ready() { return false; }
        std::coroutine_handle<> await_suspend(std::coroutine_handle<> h) {
            c->current = h;
            c->pending = true;
            return std::noop_coroutine();
        }
        HState await_resume() {
            c->pending = false;
            c->leapfrogs++;
            return std::move(c->res);
        }
    };
    Step step(const HState& s, double eps, double delta) {
        req = s;
        req_eps = eps;
        req_delta = molifier ? delta : 1e-300;
        return Step{this};
    }

    Task<double> find_epsilon() {
        const double log_target = std::log(target);
        std::vector<double> zero(dim, 0.0);
        HState s0{x, randn(), zero, 0.0};
        HState s = co_await step(s0, 0.0, 2.0);
        double eps = 1.0;
        HState x1{s.x, randn(), s.g, s.lf};
        double H0 = -x1.neg_energy();
        HState s2 = co_await step(x1, eps, 2.0 * eps);
        double prob = H0 + s2.neg_energy();
        const int direction = prob > log_target ? 1 : -1;
        for (int it = 0; it < 60; ++it) {
            if (!(direction == 1 ? prob > log_target : prob < log_target)) break;
            eps = direction == 1 ? 2.0 * eps : 0.5 * eps;
            x1.r = randn();
            H0 = -x1.neg_energy();
            s2 = co_await step(x1, eps, 2.0 * eps);
            const double p = H0 + s2.neg_energy();
            prob = std::isfinite(p) ? p : -INFINITY;
        }
        co_return eps;
    }

    Task<SubTree> buildtree(HState s, int direction, int j, double eps, double delta, double logp0,
                            double lu, Meta* meta) {
        if (j == 0) {
            HState s1 = co_await step(s, direction * eps, delta);
            const double logpprime = -s1.neg_energy();
            double lsw = logp0 - logpprime;
            if (!std::isfinite(lsw)) lsw = -INFINITY;
            const bool sprime = (lu + (logpprime - logp0)) < 1000.0;
            const double a =
                std::isfinite(logpprime) ? std::exp(std::min(0.0, logp0 - logpprime)) : 0.0;
            meta->alpha += std::isnan(a) ? 0.0 : a;
            meta->nalpha += 1;
            const int nprime = (logp0 + lu) < logpprime ? 1 : 0;
            SubTree t;
            t.xprime = s1;
            t.far = s1;
            t.near = std::move(s1);
            t.n = nprime;
            t.s = sprime;
            t.lsw = lsw;
            co_return t;
        }
        SubTree a = co_await buildtree(std::move(s), direction, j - 1, eps, delta, logp0, lu, meta);
        if (!a.s) co_return a;
        SubTree b = co_await buildtree(a.far, direction, j - 1, eps, delta, logp0, lu, meta);
        const double ls_final = logaddexp(a.lsw, b.lsw);
        if (b.lsw > ls_final || (std::isfinite(b.lsw) && U() < std::exp(b.lsw - ls_final)))
            a.xprime = std::move(b.xprime);
        const bool ok = direction == 1 ? no_uturn(a.near, b.far) : no_uturn(b.far, a.near);
        SubTree t;
        t.xprime = std::move(a.xprime);
        t.far = std::move(b.far);
        t.near = std::move(a.near);
        t.n = a.n + b.n;
        t.s = b.s && ok;
        t.lsw = ls_final;
        co_return t;
    }

    Task<int> iteration(bool adapt) {
        if (!std::isfinite(epsilon)) epsilon = co_await find_epsilon();
        if (adapt && !adapting) {
            m = 0;
            mu = std::log(10.0 * epsilon);
        }
        adapting = adapt;
        double eps;
        if (adapt) {
            m += 1;
            eps = epsilon;
        } else {
            if (m > 0) epsilon = std::exp(x_bar);
            eps = epsilon;
        }
        const double delta = 2.0 * epsilon;
        std::vector<double> zero(dim, 0.0);
        HState init{x, randn(), zero, 0.0};
        HState s0 = co_await step(init, 0.0, delta);
        HState minus = s0, plus = s0, chosen = s0;
        const double logp0 = -s0.neg_energy();
        const double lu = std::log(U());
        double log_sum_weight = 0.0;
        Meta meta;
        int j = 0;
        while (j < tree_depth) {
            const int direction = U() > 0.5 ? -1 : 1;
            meta.alpha = 0.0;
            meta.nalpha = 0;
            SubTree t;
            if (direction == -1) {
                t = co_await buildtree(minus, -1, j, eps, delta, logp0, lu, &meta);
                minus = t.far;
            } else {
                t = co_await buildtree(plus, 1, j, eps, delta, logp0, lu, &meta);
                plus = t.far;
            }
            metro_acc_prob = meta.nalpha ? std::min(1.0, meta.alpha / meta.nalpha) : 0.0;
            if (!t.s) break;
            if (t.lsw > log_sum_weight ||
                (std::isfinite(t.lsw) && U() < std::exp(t.lsw - log_sum_weight)))
                chosen = std::move(t.xprime);
            log_sum_weight = logaddexp(log_sum_weight, t.lsw);
            j += 1;
            if (!no_uturn(minus, plus)) break;
        }
        x = chosen.x;
        lf = chosen.lf;
        depth = j;
        if (adapt) {
            const double HT = target - metro_acc_prob;
            const double eta = 1.0 / (m + t0);
            s_bar = (1.0 - eta) * s_bar + eta * HT;
            const double xx = mu - s_bar * std::sqrt((double)m) / gamma;
            const double xe = std::pow((double)m, -kappa);
            x_bar = (1.0 - xe) * x_bar + xe * xx;
            epsilon = std::exp(xx);
        }
        co_return j;
    }

    // top level: n_iter transitions, recording each draw
    Task<int> run(int n_iter, bool adapt, double* draws, size_t draw_stride, double* lps,
                  int lp_stride, int chain_index) {
        for (int it = 0; it < n_iter; ++it) {
            co_await iteration(adapt);
            if (draws) std::memcpy(draws + it * draw_stride + off, x.data(), dim * sizeof(double));
            if (lps) lps[it * lp_stride + chain_index] = lf;
        }
        co_return n_iter;
    }
};

}  // namespace

struct mcp_nuts {
    std::vector<std::unique_ptr<Chain>> chains;
    size_t total = 0;
    int kind = 0;  // 1 batch, 2 mlp, 3 callback
    mcp_batch* batch = nullptr;
    mcp_mlp* mlp = nullptr;
    double prior_scale = -1.0;
    mcp_leapfrog_fn fn = nullptr;
    void* ctx = nullptr;
    std::vector<double> bx, br, bg, bed, blp;
    long long rounds = 0;
};

namespace {

int call_leapfrog(mcp_nuts* s) {
    const int n = (int)s->chains.size();
    switch (s->kind) {
        case 1: return mcp_hmc_leapfrog(s->batch, s->bed.data(), s->bx.data(), s->br.data(),
                                        s->bg.data(), s->blp.data(), nullptr);
        case 2: return mcp_mlp_leapfrog(s->mlp, n, s->prior_scale, s->bed.data(), s->bx.data(),
                                        s->br.data(), s->bg.data(), s->blp.data());
        case 3: {
            const int rc = s->fn(s->ctx, n, s->bed.data(), s->bx.data(), s->br.data(), s->bg.data(),
                                 s->blp.data());
            return rc ? set_last_error(MCP_ERR_INVALID, "leapfrog callback failed") : MCP_OK;
        }
        default: return set_last_error(MCP_ERR_STATE, "no leapfrog provider set");
    }
}

}  // namespace

extern "C" {

int mcp_nuts_create(mcp_nuts** out, int32_t n_chains, const int32_t* dims, const double* x0,
                    const uint64_t* seeds, double target, int32_t tree_depth, int32_t molifier) {
    if (!out || !dims || !x0 || !seeds) return set_last_error(MCP_ERR_INVALID, "null argument");
    *out = nullptr;
    if (n_chains < 1 || tree_depth < 1 || tree_depth > 30 || !(target > 0.0 && target < 1.0))
        return set_last_error(MCP_ERR_INVALID, "bad sampler settings");
    std::unique_ptr<mcp_nuts> s(new (std::nothrow) mcp_nuts());
    if (!s) return set_last_error(MCP_ERR_NOMEM, "host allocation failed");
    size_t off = 0;
    for (int i = 0; i < n_chains; ++i) {
        if (dims[i] < 1) return set_last_error(MCP_ERR_INVALID, "chain dimension must be >= 1");
        auto c = std::make_unique<Chain>();
        c->dim = dims[i];
        c->off = off;
        c->x.assign(x0 + off, x0 + off + dims[i]);
        for (double v : c->x)
            if (!(v > 0.0) || !std::isfinite(v))
                return set_last_error(MCP_ERR_INVALID, "initial branch lengths must be positive");
        c->rng.seed(seeds[i]);
        c->target = target;
        c->tree_depth = tree_depth;
        c->molifier = molifier != 0;
        off += dims[i];
        s->chains.push_back(std::move(c));
    }
    s->total = off;
    s->bx.resize(off);
    s->br.resize(off);
    s->bg.resize(off);
    s->bed.resize(2 * (size_t)n_chains);
    s->blp.resize(n_chains);
    *out = s.release();
    return MCP_OK;
}

void mcp_nuts_destroy(mcp_nuts* s) { delete s; }

int mcp_nuts_use_batch(mcp_nuts* s, mcp_batch* batch) {
    if (!s || !batch) return set_last_error(MCP_ERR_INVALID, "null argument");
    s->kind = 1;
    s->batch = batch;
    return MCP_OK;
}

int mcp_nuts_use_mlp(mcp_nuts* s, mcp_mlp* mlp, double prior_scale) {
    if (!s || !mlp) return set_last_error(MCP_ERR_INVALID, "null argument");
    for (auto& c : s->chains)
        if (c->dim != s->chains[0]->dim)
            return set_last_error(MCP_ERR_DIMENSION, "surrogate leapfrog needs equal chain sizes");
    s->kind = 2;
    s->mlp = mlp;
    s->prior_scale = prior_scale;
    return MCP_OK;
}

int mcp_nuts_use_callback(mcp_nuts* s, mcp_leapfrog_fn fn, void* ctx) {
    if (!s || !fn) return set_last_error(MCP_ERR_INVALID, "null argument");
    s->kind = 3;
    s->fn = fn;
    s->ctx = ctx;
    return MCP_OK;
}

int mcp_nuts_run(mcp_nuts* s, int32_t n_iter, int32_t adapt, double* draws, double* logpost,
                 int64_t* stats) {
    if (!s) return set_last_error(MCP_ERR_INVALID, "null handle");
    if (n_iter < 0) return set_last_error(MCP_ERR_INVALID, "n_iter must be >= 0");
    if (s->kind == 0) return set_last_error(MCP_ERR_STATE, "no leapfrog provider set");
    const int n = (int)s->chains.size();
    std::vector<Task<int>> tops;
    tops.reserve(n);
    for (int i = 0; i < n; ++i) {
        tops.push_back(s->chains[i]->run(n_iter, adapt != 0, draws, s->total, logpost, n, i));
        s->chains[i]->current = tops.back().h;
        s->chains[i]->pending = false;
    }
    long long rounds = 0, lf0 = 0;
    for (auto& c : s->chains) lf0 += c->leapfrogs;
    // start every chain up to its first request
    for (int i = 0; i < n; ++i) s->chains[i]->current.resume();
    while (true) {
        bool any = false;
        for (int i = 0; i < n; ++i) {
            Chain& c = *s->chains[i];
            double* x = s->bx.data() + c.off;
            double* r = s->br.data() + c.off;
            double* g = s->bg.data() + c.off;
            if (!tops[i].h.done() && c.pending) {
                any = true;
                std::memcpy(x, c.req.x.data(), c.dim * sizeof(double));
                std::memcpy(r, c.req.r.data(), c.dim * sizeof(double));
                std::memcpy(g, c.req.g.data(), c.dim * sizeof(double));
                s->bed[2 * i] = c.req_eps;
                s->bed[2 * i + 1] = c.req_delta;
            } else {
                // finished chain: a zero-length step at its position
                std::memcpy(x, c.x.data(), c.dim * sizeof(double));
                std::fill(r, r + c.dim, 0.0);
                std::fill(g, g + c.dim, 0.0);
                s->bed[2 * i] = 0.0;
                s->bed[2 * i + 1] = c.molifier ? 2.0 * (std::isfinite(c.epsilon) ? c.epsilon : 1.0)
                                               : 1e-300;
            }
        }
        if (!any) break;
        int rc = call_leapfrog(s);
        if (rc) return rc;
        ++rounds;
        for (int i = 0; i < n; ++i) {
            Chain& c = *s->chains[i];
            if (tops[i].h.done() || !c.pending) continue;
            c.res.x.assign(s->bx.begin() + c.off, s->bx.begin() + c.off + c.dim);
            c.res.r.assign(s->br.begin() + c.off, s->br.begin() + c.off + c.dim);
            c.res.g.assign(s->bg.begin() + c.off, s->bg.begin() + c.off + c.dim);
            c.res.lf = s->blp[i];
            c.current.resume();
        }
    }
    s->rounds += rounds;
    if (stats) {
        long long lf1 = 0;
        for (auto& c : s->chains) lf1 += c->leapfrogs;
        stats[0] = rounds;
        stats[1] = lf1 - lf0;
    }
    return MCP_OK;
}

int mcp_nuts_get(mcp_nuts* s, double* x, double* eps, double* lf, int32_t* depth) {
    if (!s) return set_last_error(MCP_ERR_INVALID, "null handle");
    for (size_t i = 0; i < s->chains.size(); ++i) {
        const Chain& c = *s->chains[i];
        if (x) std::memcpy(x + c.off, c.x.data(), c.dim * sizeof(double));
        if (eps) eps[i] = c.epsilon;
        if (lf) lf[i] = c.lf;
        if (depth) depth[i] = c.depth;
    }
    return MCP_OK;
}

int mcp_nuts_set(mcp_nuts* s, const double* x, const double* eps, const double* lf) {
    if (!s) return set_last_error(MCP_ERR_INVALID, "null handle");
    for (size_t i = 0; i < s->chains.size(); ++i) {
        Chain& c = *s->chains[i];
        if (x) {
            for (int k = 0; k < c.dim; ++k)
                if (!(x[c.off + k] > 0.0) || !std::isfinite(x[c.off + k]))
                    return set_last_error(MCP_ERR_INVALID, "branch lengths must be positive");
            c.x.assign(x + c.off, x + c.off + c.dim);
        }
        if (eps) c.epsilon = eps[i];
        if (lf) c.lf = lf[i];
    }
    return MCP_OK;
}

}  // extern "C"
