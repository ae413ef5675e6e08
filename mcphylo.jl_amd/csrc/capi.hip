// Host runtime and C ABI of the pruning engine.
//
// The engine owns, per batch: the datasets (tip state codes + code table +
// pattern weights), the items (topology tables, level schedules, models),
// one HIP stream, all device memory, pinned staging buffers and optional
// HIP graphs of the evaluation.  Everything topology-derived is rebuilt on
// the host (plan builder below) and uploaded; evaluation then is
//   H2D branch lengths -> k_build_P_tips -> post-order levels ->
//   [pre-order levels] -> k_reduce -> D2H (ll, grad)
// on the engine's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <queue>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "engine.h"
#include "mcphylo_gpu.h"

namespace mcp {
hipError_t launch_build_P_tips(const ModelDev*, const int32_t*, const int32_t*, int, long long,
                              long long, hipStream_t);
hipError_t launch_level(bool, int, int, const LevelJob*, const int32_t*, const int32_t*, int, int,
                        int, double*, hipStream_t);
hipError_t launch_fused(bool, int, int, const ItemSched*, int, int, int, size_t, const LevelJob*,
                        const int32_t*, const LevelJob*, const int32_t*, double*, hipStream_t);
int default_cpt(int S);
bool cpt_available(int S, int cpt);
int max_tile_cols(int S);
int tile_cols(int cpt);
hipError_t launch_reduce(const RedOp*, int, const double*, double*, double*, hipStream_t);
bool supported_states(int S);
bool mg_available(int S, int cpt);
hipError_t launch_level_mg(int, int, const LevelJob*, const int32_t*, const int32_t*, int, int, int,
                           double*, double*, hipStream_t);
hipError_t launch_lf_pre(const HmcItem*, int, double*, double*, const double*, const double*,
                         double*, int, hipStream_t);
hipError_t launch_lf_post(const HmcItem*, int, const double*, double*, double*, const double*,
                          const double*, double*, const uint8_t*, hipStream_t);
}  // namespace mcp

using namespace mcp;

// Grid of the level kernels: at most this many single-wave blocks per compute
// unit, beyond which a block takes several consecutive tasks (1024 measured
// 3% faster than 256 and 6% faster than 64 on the 256-tree ragged batch).
#ifndef MCP_GRID_PER_CU
#define MCP_GRID_PER_CU 1024
#endif

namespace {

thread_local std::string g_last_error;

struct Status {
    int code;
    std::string msg;
};

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

}  // namespace

namespace mcp {
// error reporting for the other translation units of the library
int set_last_error(int code, const char* msg) { return fail(code, msg); }
}  // namespace mcp

namespace {

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            return fail(e_ == hipErrorOutOfMemory ? MCP_ERR_NOMEM : MCP_ERR_HIP,            \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                 \
        }                                                                                   \
    } while (0)

size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

// Device buffer with capacity (bytes); grows by reallocation.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int ensure(size_t bytes) {
        if (bytes <= cap && p) return MCP_OK;
        if (p) {
            (void)hipFree(p);
            p = nullptr;
            cap = 0;
        }
        size_t b = std::max<size_t>(bytes, 256);
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(MCP_ERR_NOMEM, std::string("hipMalloc of ") + std::to_string(b) +
                                           " bytes failed: " + hipGetErrorString(e));
        }
        cap = b;
        return MCP_OK;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

struct PinnedBuf {
    void* p = nullptr;
    size_t cap = 0;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    int ensure(size_t bytes) {
        if (bytes <= cap && p) return MCP_OK;
        if (p) {
            (void)hipHostFree(p);
            p = nullptr;
            cap = 0;
        }
        size_t b = std::max<size_t>(bytes, 256);
        hipError_t e = hipHostMalloc(&p, b, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(MCP_ERR_NOMEM, "hipHostMalloc failed");
        }
        cap = b;
        return MCP_OK;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Host-side record of one node of an item (index = num-1).
struct NodeInfo {
    int32_t first_child;   // offset into ItemHost::children
    int32_t nchild;
    int32_t fslot;         // partial slab index (internal non-root), else -1
    int32_t leaf;          // row of the dataset's tip codes (leaves), else -1
    int32_t gslot;         // pre-order slab index in ring buffer gbuf (internal non-root), else -1
    int32_t gbuf;          // 0/1: which pre-order ring buffer (depth parity)
    int32_t fl;            // fused schedule: slot of the partial's post-order copy in the
                           // block's LDS slab stage (internal non-root), else -1
    int32_t lpart;         // offset (in the item) of the log-scaler partial sums (internal), else -1
    int32_t gpart;         // offset (in the item) of the gradient partial sums (non-root), else -1
    int32_t is_root;
};

// Largest tip-code stage (bytes of LDS per block) of a fused launch: at 64
// patterns per tile, 512 leaves.  Larger items read their codes from global
// memory.
constexpr size_t kLdsCodesMax = 32 * 1024;

// Largest LDS stage of a fused block (tip codes, then the pre-order slabs of
// the fused walk): three such blocks fit one CU's 160 KB.  (The 100-taxon,
// 4-state tree at 64 patterns per tile needs 6.4 KB of codes and 22 slabs of
// 2 KB.)
constexpr size_t kLdsFusedMax = 52 * 1024;

struct DatasetHost {
    int S = 0, C = 0, Cp = 0, K = 0, nleaves = 0;
    std::vector<int32_t> leaf_nums;          // row j of codes is leaf leaf_nums[j]
    std::unordered_map<int32_t, int32_t> row_of;
    std::vector<uint8_t> codes;              // [nleaves][Cp]
    std::vector<double> tiptab;              // [K][S]
    std::vector<double> weights;             // [Cp]
    size_t codes_off = 0, tiptab_off = 0, weights_off = 0;  // offsets in device arenas
};

struct ItemHost {
    int dataset = 0, N = 0, R = 0, S = 0, nint = 0;
    std::vector<int32_t> parent;             // [N]
    std::vector<int32_t> child_ptr, children;  // CSR by node index (num-1), children ascending
    std::vector<NodeInfo> nodes;
    std::vector<int32_t> height, depth;
    int max_height = 0, max_depth = 0;
    int g_slots[2] = {0, 0};
    int fl_slots = 0;        // LDS slots of the post-order frontier (fused schedule)
    bool lds_stage = false;  // fused walk keeps its frontier slabs in LDS
    // model
    bool model_set = false;
    std::vector<double> U, D, Uinv, rates, pi;
    double mu = 0.0;
    // layout
    size_t F_off = 0;      // doubles into F arena
    size_t G_off[2] = {0, 0};  // doubles into G arenas
    size_t P_off = 0;      // doubles into P/dP arenas ((N-1)*R*S*S)
    size_t PT_off = 0;     // doubles into PT/dPT arenas (nleaves*R*K*S)
    size_t model_off = 0;  // doubles into model arena
    size_t blv_off = 0;    // doubles into blv (N-1)
    int parts_off = 0;
    int out_off = 0;
    int ntiles = 0;
    long long mg_off = -1;  // doubles into the model-gradient partial sums, -1 if not laid out
    int children_off = 0;  // int32 into children arena
    int nodes_off = 0;     // NodeInfo into nodes arena
    int leafof_off = 0;    // int32 into leaf_of arena
};

struct Launch {
    bool adjoint;
    int S, cpt;
    double bytes;       // algorithmic HBM bytes of the launch
    int job_begin, njobs;
    int taskoff_begin;  // index into task_off arena (njobs+1 entries)
    int ntasks;
    int grid, chunk;    // blocks and tasks per block
    int blk_begin;      // index into the block -> first job arena (grid entries)
};

}  // namespace

struct mcp_engine {
    int device = 0;
    int waves = 4;     // wavefronts per block of the fused kernel (1..8)
    int schedule = MCP_SCHEDULE_LEVELS;
    int hybrid_width = 4;  // widest level (internal nodes) run fused under HYBRID
    // under the fused schedule, keep an item's pre-order slabs in the block's
    // LDS when they fit (MCP_FUSED_LDS_SLABS=0 at engine creation: global rings)
    bool lds_slabs = true;
    int num_cu = 256;
    hipStream_t stream = nullptr;
    std::vector<DatasetHost> ds;
    std::vector<ItemHost> items;
    // device arenas
    DevBuf dF, dG0, dG1, dP, ddP, dPT, ddPT, dparts, dout, dblv, dmodel, dcodes, dweights, dtiptab,
        dleafof, dmodels, djobs_f, djobs_a, dmore, dtaskoff, dred, delemP, delemT, dsched, dlvl_f,
        dlvl_a, dblkjob, dred2, dredtmp;
    PinnedBuf hblv, hout;
    // schedule
    std::vector<Launch> fwd, adj;
    std::vector<int32_t> taskoff, blkjob;
    struct FusedLaunch {
        int S, cpt, sched_begin, nitems, nblocks;
        bool tail;  // deep narrow pre-order levels, run after the per-level launches
        size_t lds_bytes;  // LDS stage per block (max over the launch's items)
        double bytes_f, bytes_a;  // algorithmic bytes of its post-/pre-order levels
    };
    // columns per thread chosen for each state count (-1: the built default);
    // set by mcp_set_tiles / mcp_autotune
    int cpt_sel[33] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1,
                       -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
    int cpt(int S) const { return (S < 33 && cpt_sel[S] >= 0) ? cpt_sel[S] : default_cpt(S); }
    bool evaluated = false;  // branch lengths uploaded at least once
    std::vector<FusedLaunch> fused;
    // reductions: stage 1 over the partial sums (ll ops first, then gradient
    // ops; long ones split into chunks into a temporary), stage 2 combining
    // the chunks (ll ops first)
    std::vector<RedOp> red, red2;
    int n_ll_ops = 0, n_ll_ops2 = 0;
    // per-(branch, category) gradient components: one sum over the tiles of
    // each (item, branch, category) into dout_vr (layout [item][branch][r])
    std::vector<RedOp> red_vr;
    std::vector<size_t> vr_off;
    size_t total_vr = 0;
    DevBuf dred_vr, dout_vr;
    PinnedBuf hout_vr;
    // model gradient (d logL / dP per branch and category, d logL / d pi):
    // every pre-order level as k_level_mg launches (whatever the schedule),
    // partial sums in dmparts (allocated on first use), one sum over the tiles
    // per output into dout_mg (layout per item: [branch][r][S*S], then [S])
    std::vector<Launch> adj_mg;
    bool mg_ok = false;  // every item has model-gradient kernels and 32-bit offsets fit
    long long total_mparts = 0;
    std::vector<RedOp> red_mg;
    std::vector<size_t> mg_out_off;
    size_t total_mg = 0;
    DevBuf dmparts, dred_mg, dout_mg;
    PinnedBuf hout_mg;
    long long nelemP = 0, nelemT = 0;
    size_t total_blv = 0, total_out = 0;
    long long total_patterns = 0;
    double bytes_fwd = 0.0, bytes_adj = 0.0;
    // graphs
    bool use_graph = false;
    hipGraphExec_t graph_exec[3] = {nullptr, nullptr, nullptr};  // logl, logl+grad, leapfrog
    // device leapfrog: [x | r | g] per item (blv layout), eps/delta pairs, log posterior
    DevBuf dhmc, dhmc_items, dlp, dhmc_int;  // dhmc_int: 1 per internal branch (blv layout)
    PinnedBuf hhmc;
    std::vector<HmcItem> hmc_items;
    bool hmc_items_dirty = true;
    // timing: one event pair around every kernel launch of an evaluation
    bool timing = false;
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    std::vector<hipEvent_t> kev;  // [2k] before launch k, [2k+1] after
    std::vector<int> kphase;      // phase of launch k: 1 P build, 2 post-order, 3 pre-order, 4 reduce
    std::vector<double> kbytes;   // algorithmic bytes of launch k (0 outside the level passes)
    std::vector<double> kms;      // duration of launch k in the last timed evaluation
    int nk = 0;
    double last_ms[5] = {0, 0, 0, 0, 0};

    ~mcp_engine() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        for (auto& e : kev)
            if (e) (void)hipEventDestroy(e);
        if (ev_begin) (void)hipEventDestroy(ev_begin);
        if (ev_end) (void)hipEventDestroy(ev_end);
        if (stream) (void)hipStreamDestroy(stream);
    }
    void drop_graphs() {
        for (auto& g : graph_exec) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    }
};

namespace {

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// ---------------------------------------------------------------------------
// datasets: tip partials -> state codes (+ optional pattern compression)
// ---------------------------------------------------------------------------

int build_dataset(const mcp_dataset_desc& d, DatasetHost& out) {
    if (d.n_states < 1 || d.n_patterns < 1 || d.n_leaves < 1 || !d.leaf_nums ||
        (!d.tip_partials && (!d.tip_codes || !d.code_table || d.n_codes < 1)))
        return fail(MCP_ERR_INVALID, "dataset: sizes must be positive and pointers non-null");
    if (!d.tip_partials && d.n_codes > 256)
        return fail(MCP_ERR_UNSUPPORTED, "dataset: more than 256 tip codes");
    if (!supported_states(d.n_states))
        return fail(MCP_ERR_UNSUPPORTED, "dataset: no kernel built for " +
                                             std::to_string(d.n_states) + " states");
    const int S = d.n_states, C = d.n_patterns, L = d.n_leaves;
    out.S = S;
    out.nleaves = L;
    out.leaf_nums.assign(d.leaf_nums, d.leaf_nums + L);
    // sort leaves by num so rows are canonical
    std::vector<int> order(L);
    for (int j = 0; j < L; ++j) order[j] = j;
    std::sort(order.begin(), order.end(),
              [&](int a, int b) { return d.leaf_nums[a] < d.leaf_nums[b]; });
    std::vector<int32_t> sorted_nums(L);
    for (int j = 0; j < L; ++j) sorted_nums[j] = d.leaf_nums[order[j]];
    for (int j = 1; j < L; ++j)
        if (sorted_nums[j] == sorted_nums[j - 1])
            return fail(MCP_ERR_INVALID, "dataset: duplicate leaf num");
    out.leaf_nums = sorted_nums;
    out.row_of.clear();
    for (int j = 0; j < L; ++j) out.row_of[sorted_nums[j]] = j;

    // code table over distinct tip vectors
    std::unordered_map<std::string, int> code_of;
    std::vector<uint8_t> codes((size_t)L * C);
    out.tiptab.clear();
    if (!d.tip_partials) {
        for (int j = 0; j < L; ++j) {
            const uint8_t* src = d.tip_codes + (size_t)order[j] * C;
            for (int c = 0; c < C; ++c) {
                if (src[c] >= d.n_codes)
                    return fail(MCP_ERR_INVALID, "dataset: tip code out of range");
                codes[(size_t)j * C + c] = src[c];
            }
        }
        out.tiptab.assign(d.code_table, d.code_table + (size_t)d.n_codes * S);
        out.K = d.n_codes;
    }
    for (int j = 0; j < L && d.tip_partials; ++j) {
        const double* xl = d.tip_partials + (size_t)S * C * order[j];
        for (int c = 0; c < C; ++c) {
            std::string key(reinterpret_cast<const char*>(xl + (size_t)S * c), sizeof(double) * S);
            auto it = code_of.find(key);
            int code;
            if (it == code_of.end()) {
                code = (int)code_of.size();
                if (code > 255)
                    return fail(MCP_ERR_UNSUPPORTED,
                                "dataset: more than 256 distinct tip vectors");
                code_of.emplace(key, code);
                out.tiptab.insert(out.tiptab.end(), xl + (size_t)S * c, xl + (size_t)S * (c + 1));
            } else {
                code = it->second;
            }
            codes[(size_t)j * C + c] = (uint8_t)code;
        }
    }
    if (d.tip_partials) out.K = (int)code_of.size();

    std::vector<double> w(C, 1.0);
    if (d.pattern_weights) w.assign(d.pattern_weights, d.pattern_weights + C);
    int Cu = C;
    std::vector<int> colmap(C);
    for (int c = 0; c < C; ++c) colmap[c] = c;
    if (d.flags & MCP_FLAG_COMPRESS_PATTERNS) {
        std::unordered_map<std::string, int> col_of;
        std::vector<double> wu;
        std::vector<int> keep;
        std::string key(L, '\0');
        for (int c = 0; c < C; ++c) {
            for (int j = 0; j < L; ++j) key[j] = (char)codes[(size_t)j * C + c];
            auto it = col_of.find(key);
            if (it == col_of.end()) {
                col_of.emplace(key, (int)keep.size());
                keep.push_back(c);
                wu.push_back(w[c]);
            } else {
                wu[it->second] += w[c];
            }
        }
        Cu = (int)keep.size();
        colmap = keep;
        w = wu;
    }
    out.C = Cu;
    out.Cp = (int)round_up((size_t)Cu, (size_t)max_tile_cols(S));
    out.codes.assign((size_t)L * out.Cp, 0);
    for (int j = 0; j < L; ++j)
        for (int c = 0; c < Cu; ++c)
            out.codes[(size_t)j * out.Cp + c] = codes[(size_t)j * C + colmap[c]];
    out.weights.assign(out.Cp, 0.0);
    for (int c = 0; c < Cu; ++c) out.weights[c] = w[c];
    return MCP_OK;
}

// ---------------------------------------------------------------------------
// items: topology tables
// ---------------------------------------------------------------------------

void assign_partial_slots(ItemHost& it);

// Pre-order slabs of a level walk: the slab of node v (depth d) is written
// at level d - 1 and read at level d, so two ring buffers by depth parity,
// with slots reused from depth d to depth d + 2, hold them.
void assign_g_slots_by_depth(ItemHost& it) {
    std::vector<int> gcount(it.max_depth + 2, 0);
    for (int v = 0; v < it.N; ++v) {
        NodeInfo& n = it.nodes[v];
        n.gbuf = it.depth[v] & 1;
        n.gslot = (n.nchild > 0 && !n.is_root) ? gcount[it.depth[v]]++ : -1;
    }
    it.g_slots[0] = it.g_slots[1] = 0;
    for (int dd = 0; dd <= it.max_depth; ++dd)
        it.g_slots[dd & 1] = std::max(it.g_slots[dd & 1], gcount[dd]);
}

// The same for a walk in arbitrary pre-order steps (step[v]: the step of
// node v's pre-order task): the slab of v lives from its mother's step
// (written) to its own (read).  Interval colouring in one buffer; a slot is
// reused only by a slab written in a later step than its last reader's.
void assign_g_slots_by_steps(ItemHost& it, const std::vector<int>& step) {
    std::vector<int> order;
    for (int v = 0; v < it.N; ++v)
        if (it.nodes[v].nchild > 0 && !it.nodes[v].is_root) order.push_back(v);
    auto start = [&](int v) { return step[it.parent[v] - 1]; };
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return start(a) < start(b); });
    std::priority_queue<std::pair<int, int>, std::vector<std::pair<int, int>>,
                        std::greater<std::pair<int, int>>>
        busy;  // (last step read, slot)
    std::priority_queue<int, std::vector<int>, std::greater<int>> free_slots;
    int nslots = 0;
    for (int v = 0; v < it.N; ++v) {
        it.nodes[v].gbuf = 0;
        it.nodes[v].gslot = -1;
    }
    for (int v : order) {
        while (!busy.empty() && busy.top().first < start(v)) {
            free_slots.push(busy.top().second);
            busy.pop();
        }
        int slot;
        if (!free_slots.empty()) {
            slot = free_slots.top();
            free_slots.pop();
        } else {
            slot = nslots++;
        }
        it.nodes[v].gslot = slot;
        busy.push({step[v], slot});
    }
    it.g_slots[0] = nslots;
    it.g_slots[1] = 0;
}

// Post-order frontier of a walk in steps (step[v]: the step of node v's
// post-order task): the partial of v is read by its mother's task, so a copy
// lives from step[v] (written) to its mother's step (read).  Interval
// colouring as above; slots are NodeInfo::fl.
void assign_fl_slots_by_steps(ItemHost& it, const std::vector<int>& step) {
    std::vector<int> order;
    for (int v = 0; v < it.N; ++v) {
        it.nodes[v].fl = -1;
        if (it.nodes[v].nchild > 0 && !it.nodes[v].is_root) order.push_back(v);
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return step[a] < step[b]; });
    std::priority_queue<std::pair<int, int>, std::vector<std::pair<int, int>>,
                        std::greater<std::pair<int, int>>>
        busy;  // (step read, slot)
    std::priority_queue<int, std::vector<int>, std::greater<int>> free_slots;
    int nslots = 0;
    for (int v : order) {
        while (!busy.empty() && busy.top().first < step[v]) {
            free_slots.push(busy.top().second);
            busy.pop();
        }
        int slot;
        if (!free_slots.empty()) {
            slot = free_slots.top();
            free_slots.pop();
        } else {
            slot = nslots++;
        }
        it.nodes[v].fl = slot;
        busy.push({step[it.parent[v] - 1], slot});
    }
    it.fl_slots = nslots;
}

int build_topology(ItemHost& it, const DatasetHost& d, const int32_t* parent) {
    const int N = it.N;
    if (N < 2) return fail(MCP_ERR_INVALID, "tree needs at least two nodes");
    it.parent.assign(parent, parent + N);
    if (it.parent[N - 1] != 0) return fail(MCP_ERR_INVALID, "the root must be node num N");
    std::vector<int32_t> cnt(N, 0);
    for (int v = 0; v < N - 1; ++v) {
        int p = it.parent[v];
        if (p < 1 || p > N || p == v + 1)
            return fail(MCP_ERR_INVALID, "parent num out of range at node " + std::to_string(v + 1));
        cnt[p - 1]++;
    }
    it.child_ptr.assign(N + 1, 0);
    for (int v = 0; v < N; ++v) it.child_ptr[v + 1] = it.child_ptr[v] + cnt[v];
    it.children.assign(N - 1, 0);
    std::vector<int32_t> fill(it.child_ptr.begin(), it.child_ptr.end() - 1);
    for (int v = 0; v < N - 1; ++v) it.children[fill[it.parent[v] - 1]++] = v;
    // depth by walking up (detects cycles)
    it.depth.assign(N, -1);
    it.depth[N - 1] = 0;
    for (int v = 0; v < N; ++v) {
        std::vector<int> path;
        int x = v, steps = 0;
        while (it.depth[x] < 0) {
            path.push_back(x);
            x = it.parent[x] - 1;
            if (++steps > N) return fail(MCP_ERR_INVALID, "parent array has a cycle");
        }
        int dd = it.depth[x];
        for (int k = (int)path.size() - 1; k >= 0; --k) it.depth[path[k]] = ++dd;
    }
    // leaves must match the dataset
    int nleaves = 0;
    for (int v = 0; v < N; ++v)
        if (cnt[v] == 0) {
            ++nleaves;
            if (!d.row_of.count(v + 1))
                return fail(MCP_ERR_DIMENSION,
                            "leaf num " + std::to_string(v + 1) + " has no tip data");
        }
    if (nleaves != d.nleaves)
        return fail(MCP_ERR_DIMENSION, "tree leaves and tip data disagree");
    if (cnt[N - 1] == 0) return fail(MCP_ERR_INVALID, "root has no children");
    // heights (post-order levels): process nodes by decreasing depth
    std::vector<int> by_depth(N);
    for (int v = 0; v < N; ++v) by_depth[v] = v;
    std::stable_sort(by_depth.begin(), by_depth.end(),
                     [&](int a, int b) { return it.depth[a] > it.depth[b]; });
    it.height.assign(N, 0);
    for (int v : by_depth) {
        int h = 0;
        for (int q = it.child_ptr[v]; q < it.child_ptr[v + 1]; ++q)
            h = std::max(h, it.height[it.children[q]] + 1);
        it.height[v] = h;
    }
    it.max_height = it.height[N - 1];
    it.max_depth = *std::max_element(it.depth.begin(), it.depth.end());
    // node records
    it.nodes.assign(N, NodeInfo{});
    it.nint = 0;
    int fslot = 0, lidx = 0;
    std::vector<int> gcount(it.max_depth + 2, 0);
    for (int v = 0; v < N; ++v) {
        NodeInfo& n = it.nodes[v];
        n.first_child = it.child_ptr[v];
        n.nchild = cnt[v];
        n.is_root = (v == N - 1);
        n.fslot = -1;
        n.leaf = -1;
        n.gslot = -1;
        n.gbuf = it.depth[v] & 1;
        n.lpart = -1;
        n.gpart = -1;
        if (cnt[v] == 0) {
            n.leaf = d.row_of.at(v + 1);
        } else {
            it.nint++;
            if (!n.is_root) n.fslot = fslot++;
        }
    }
    assign_g_slots_by_depth(it);
    (void)gcount;
    (void)fslot;
    (void)lidx;
    assign_partial_slots(it);
    return MCP_OK;
}

// Partial-sum slots of every node relative to the item (made absolute in the
// layout): R * ntiles per internal node (log scalers), then per branch
// (gradient); they follow the item's tile count.
void assign_partial_slots(ItemHost& it) {
    const int per = it.R * it.ntiles;
    int lidx = 0;
    for (int v = 0; v < it.N; ++v) {
        NodeInfo& n = it.nodes[v];
        if (n.nchild > 0) n.lpart = (lidx++) * per;
    }
    for (int v = 0; v < it.N - 1; ++v) it.nodes[v].gpart = (it.nint + v) * per;
}

int item_from_desc(const mcp_item_desc& d, const std::vector<DatasetHost>& ds, ItemHost& it) {
    if (d.dataset < 0 || d.dataset >= (int)ds.size())
        return fail(MCP_ERR_INVALID, "item: dataset index out of range");
    if (!d.parent || d.n_nodes < 2 || d.n_rates < 1)
        return fail(MCP_ERR_INVALID, "item: bad sizes or null parent array");
    const DatasetHost& D = ds[d.dataset];
    it.dataset = d.dataset;
    it.N = d.n_nodes;
    it.R = d.n_rates;
    it.S = D.S;
    it.ntiles = D.Cp / tile_cols(default_cpt(D.S));
    return build_topology(it, D, d.parent);
}

}  // namespace

// ---------------------------------------------------------------------------
// layout, schedule, upload
// ---------------------------------------------------------------------------

static int upload(mcp_engine* e, DevBuf& b, const void* src, size_t bytes) {
    int st = b.ensure(bytes);
    if (st) return st;
    if (bytes) HIP_TRY(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    (void)e;
    return MCP_OK;
}

// Device descriptors and launch schedule; called after create and after any
// topology change.  F/P/model arenas keep their layout (sizes do not depend
// on topology); the pre-order ring buffers may grow, so every job record is
// rebuilt with fresh addresses.
// Hu's list schedule of an item's internal nodes on W processors (the
// wavefronts of a fused block): steps of at most W nodes whose internal
// children lie in earlier steps, the ready nodes deepest first (ties: lower
// index).  For unit-time tasks whose precedence graph is an in-tree (a
// post-order pass) this minimises the number of steps; run backwards it is
// a schedule of the pre-order pass (an out-tree) with the same count.  The
// 100-taxon tree needs 26 steps at W = 4 where its 14 height levels take 32
// task slots per wavefront.
static std::vector<std::vector<int>> hu_steps(const ItemHost& it, int W) {
    std::vector<int> pending(it.N, 0);
    for (int v = 0; v < it.N; ++v) {
        const int p = it.parent[v] - 1;
        if (p >= 0 && it.nodes[v].nchild > 0) pending[p]++;
    }
    auto deeper = [&](int a, int b) {
        return it.depth[a] != it.depth[b] ? it.depth[a] < it.depth[b] : a > b;
    };
    std::priority_queue<int, std::vector<int>, decltype(deeper)> ready(deeper);
    for (int v = 0; v < it.N; ++v)
        if (it.nodes[v].nchild > 0 && pending[v] == 0) ready.push(v);
    std::vector<std::vector<int>> steps;
    while (!ready.empty()) {
        std::vector<int> step;
        while (!ready.empty() && (int)step.size() < W) {
            step.push_back(ready.top());
            ready.pop();
        }
        for (int v : step) {
            const int p = it.parent[v] - 1;
            if (p >= 0 && --pending[p] == 0) ready.push(p);
        }
        steps.push_back(std::move(step));
    }
    return steps;
}

static int rebuild_schedule(mcp_engine* e) {
    e->drop_graphs();
    e->hmc_items_dirty = true;
    const int nitems = (int)e->items.size();
    // the fused schedule walks each item in Hu steps; pre-order slabs follow
    // the walk's order
    std::vector<std::vector<std::vector<int>>> hu(nitems);
    for (int i = 0; i < nitems; ++i) {
        ItemHost& it = e->items[i];
        if (e->schedule == MCP_SCHEDULE_FUSED) {
            hu[i] = hu_steps(it, e->waves);
            std::vector<int> pstep(it.N, -1), fstep(it.N, -1);
            const int ns = (int)hu[i].size();
            for (int k = 0; k < ns; ++k)
                for (int v : hu[i][k]) {
                    fstep[v] = k;
                    pstep[v] = ns - 1 - k;  // pre-order runs the steps backwards
                }
            assign_g_slots_by_steps(it, pstep);
            assign_fl_slots_by_steps(it, fstep);
        } else {
            assign_g_slots_by_depth(it);
            it.fl_slots = 0;
            for (NodeInfo& n : it.nodes) n.fl = -1;
        }
    }
    // pre-order ring buffers
    size_t g0 = 0, g1 = 0;
    for (auto& it : e->items) {
        const DatasetHost& d = e->ds[it.dataset];
        const size_t slab = (size_t)it.R * d.Cp * it.S;
        it.G_off[0] = g0;
        it.G_off[1] = g1;
        g0 += (size_t)it.g_slots[0] * slab;
        g1 += (size_t)it.g_slots[1] * slab;
    }
    int st;
    if ((st = e->dG0.ensure(std::max<size_t>(g0, 1) * sizeof(double)))) return st;
    if ((st = e->dG1.ensure(std::max<size_t>(g1, 1) * sizeof(double)))) return st;

    // P-build records and leaf maps
    std::vector<int32_t> leaf_of;
    std::vector<ModelDev> dmodels(nitems);
    std::vector<int32_t> elemP(nitems + 1, 0), elemT(nitems + 1, 0);
    e->bytes_fwd = e->bytes_adj = 0.0;
    for (int i = 0; i < nitems; ++i) {
        ItemHost& it = e->items[i];
        const DatasetHost& d = e->ds[it.dataset];
        it.leafof_off = (int)leaf_of.size();
        for (const NodeInfo& n : it.nodes) leaf_of.push_back(n.leaf);
        elemP[i + 1] = elemP[i] + (it.N - 1) * it.R * it.S * it.S;
        elemT[i + 1] = elemT[i] + (it.N - 1) * it.R * d.K * it.S;
        // algorithmic bytes
        const double w = 8.0 * it.S * it.R * d.C;
        for (int v = 0; v < it.N; ++v) {
            const NodeInfo& n = it.nodes[v];
            if (n.nchild == 0) {
                e->bytes_adj += d.C;  // leaf code read in the pre-order pass
                continue;
            }
            for (int q = it.child_ptr[v]; q < it.child_ptr[v + 1]; ++q) {
                const int u = it.children[q];
                e->bytes_fwd += it.nodes[u].nchild ? w : (double)d.C;
            }
            if (!n.is_root) {
                e->bytes_fwd += w;        // write F_v
                e->bytes_adj += w;        // read G_v as a parent
                e->bytes_adj += 2.0 * w;  // read F_v, write G_v as a child
            }
        }
    }
    if ((st = upload(e, e->dleafof, leaf_of.data(), leaf_of.size() * sizeof(int32_t)))) return st;
    for (int i = 0; i < nitems; ++i) {
        ItemHost& it = e->items[i];
        const DatasetHost& d = e->ds[it.dataset];
        ModelDev& m = dmodels[i];
        const double* mb = e->dmodel.as<double>() + it.model_off;
        m.U = mb;
        m.D = mb + it.S * it.S;
        m.Uinv = mb + it.S * it.S + it.S;
        m.rates = mb + 2 * it.S * it.S + it.S;
        m.blv = e->dblv.as<double>() + it.blv_off;
        m.tiptab = e->dtiptab.as<double>() + d.tiptab_off;
        m.leaf_of = e->dleafof.as<int32_t>() + it.leafof_off;
        m.mu = it.mu;
        m.S = it.S;
        m.R = it.R;
        m.N = it.N;
        m.K = d.K;
        m.P = e->dP.as<double>() + it.P_off;
        m.dP = e->ddP.as<double>() + it.P_off;
        m.PT = e->dPT.as<double>() + it.PT_off;
        m.dPT = e->ddPT.as<double>() + it.PT_off;
    }
    if ((st = upload(e, e->dmodels, dmodels.data(), dmodels.size() * sizeof(ModelDev)))) return st;
    if ((st = upload(e, e->delemP, elemP.data(), elemP.size() * sizeof(int32_t)))) return st;
    if ((st = upload(e, e->delemT, elemT.data(), elemT.size() * sizeof(int32_t)))) return st;
    e->nelemP = elemP[nitems];
    e->nelemT = elemT[nitems];

    // children beyond the inline ones: reserve the overflow arena first so
    // job records can point into it
    size_t nmore = 0;
    for (auto& it : e->items)
        for (const NodeInfo& n : it.nodes) nmore += std::max(0, n.nchild - kInlineChildren);
    if ((st = e->dmore.ensure(std::max<size_t>(nmore, 1) * sizeof(ChildRef)))) return st;
    std::vector<ChildRef> more;
    more.reserve(nmore);

    auto child_ref = [&](const ItemHost& it, const DatasetHost& d, int u) {
        const NodeInfo& n = it.nodes[u];
        ChildRef c{};
        const size_t slab = (size_t)it.R * d.Cp * it.S;
        if (n.nchild == 0) {
            c.F = nullptr;
            c.codes = e->dcodes.as<uint8_t>() + d.codes_off + (size_t)n.leaf * d.Cp;
            c.P = e->dPT.as<double>() + it.PT_off + (size_t)n.leaf * it.R * d.K * it.S;
            c.dP = e->ddPT.as<double>() + it.PT_off + (size_t)n.leaf * it.R * d.K * it.S;
            c.G = nullptr;
            c.leaf = n.leaf;
            c.gslot = -1;
            c.fslot = -1;
        } else {
            c.F = e->dF.as<double>() + it.F_off + (size_t)n.fslot * slab;
            c.codes = nullptr;
            c.P = e->dP.as<double>() + it.P_off + (size_t)u * it.R * it.S * it.S;
            c.dP = e->ddP.as<double>() + it.P_off + (size_t)u * it.R * it.S * it.S;
            double* gb = n.gbuf ? e->dG1.as<double>() + it.G_off[1] : e->dG0.as<double>() + it.G_off[0];
            c.G = gb + (size_t)n.gslot * slab;
            c.leaf = -1;
            c.gslot = it.lds_stage ? (int16_t)n.gslot : -1;
            c.fslot = it.lds_stage ? (int16_t)n.fl : -1;
        }
        c.gpart = it.parts_off + n.gpart;
        c.mpart = it.mg_off < 0 ? -1
                                : (int32_t)(it.mg_off + (long long)u * it.R * it.S * it.S * it.ntiles);
        return c;
    };
    auto make_job = [&](int i, int v) {
        const ItemHost& it = e->items[i];
        const DatasetHost& d = e->ds[it.dataset];
        const NodeInfo& n = it.nodes[v];
        const size_t slab = (size_t)it.R * d.Cp * it.S;
        LevelJob J{};
        J.S = it.S;
        J.C = d.C;
        J.Cp = d.Cp;
        J.R = it.R;
        J.K = d.K;
        J.ntiles = it.ntiles;
        J.nchild = n.nchild;
        J.lpart = it.parts_off + n.lpart;
        J.is_root = n.is_root;
        J.F = n.is_root ? nullptr : e->dF.as<double>() + it.F_off + (size_t)n.fslot * slab;
        J.gslot = (n.is_root || !it.lds_stage) ? -1 : n.gslot;
        J.fslot = (n.is_root || !it.lds_stage) ? -1 : n.fl;
        if (n.is_root) {
            J.G = nullptr;
        } else {
            const double* gb = n.gbuf ? e->dG1.as<double>() + it.G_off[1]
                                      : e->dG0.as<double>() + it.G_off[0];
            J.G = gb + (size_t)n.gslot * slab;
        }
        J.weights = e->dweights.as<double>() + d.weights_off;
        J.pi = e->dmodel.as<double>() + it.model_off + 2 * it.S * it.S + it.S + it.R;
        J.tiptab = e->dtiptab.as<double>() + d.tiptab_off;
        J.rpart = (n.is_root && it.mg_off >= 0)
                      ? (int32_t)(it.mg_off + (long long)(it.N - 1) * it.R * it.S * it.S * it.ntiles)
                      : -1;
        J.more = nullptr;
        for (int q = 0; q < n.nchild; ++q) {
            const ChildRef c = child_ref(it, d, it.children[n.first_child + q]);
            if (q < kInlineChildren) {
                J.kid[q] = c;
            } else {
                if (q == kInlineChildren) J.more = e->dmore.as<ChildRef>() + more.size();
                more.push_back(c);
            }
        }
        return J;
    };

    // level schedule: post-order by height, pre-order by depth, per state count
    std::vector<LevelJob> jf, ja;
    e->taskoff.clear();
    e->blkjob.clear();
    e->fwd.clear();
    e->adj.clear();
    e->fused.clear();
    int maxh = 0, maxd = 0;
    std::vector<int> svals;
    for (auto& it : e->items) {
        maxh = std::max(maxh, it.max_height);
        maxd = std::max(maxd, it.max_depth);
        if (std::find(svals.begin(), svals.end(), it.S) == svals.end()) svals.push_back(it.S);
    }
    std::sort(svals.begin(), svals.end());
    // Per item, the levels run by the fused kernel: post-order heights >= h0
    // and pre-order depths <= d0 (consecutive in time: the end of the
    // post-order pass and the start of the pre-order one).  LEVELS: none;
    // FUSED: all; HYBRID: the top levels of at most hybrid_width internal
    // nodes, so narrow levels share one launch, and a second fused launch
    // ("tail") walks the pre-order depths >= d1 when every level from d1 down
    // is that narrow too (at least two of them).
    std::vector<int> h0(nitems), d0(nitems), d1(nitems);
    std::vector<int> fused_first(nitems, -1);  // first post-order job of a fused-schedule item
    for (int i = 0; i < nitems; ++i) {
        const ItemHost& it = e->items[i];
        d1[i] = it.max_depth + 1;
        if (e->schedule == MCP_SCHEDULE_LEVELS) {
            h0[i] = it.max_height + 1;
            d0[i] = -1;
        } else if (e->schedule == MCP_SCHEDULE_FUSED) {
            h0[i] = 1;
            d0[i] = it.max_depth;
        } else {
            std::vector<int> wh(it.max_height + 1, 0), wd(it.max_depth + 1, 0);
            for (int v = 0; v < it.N; ++v)
                if (it.nodes[v].nchild > 0) {
                    wh[it.height[v]]++;
                    wd[it.depth[v]]++;
                }
            int h = it.max_height;
            while (h >= 1 && wh[h] <= e->hybrid_width) --h;
            h0[i] = h + 1;
            int dd = 0;
            while (dd <= it.max_depth && wd[dd] <= e->hybrid_width) ++dd;
            d0[i] = dd - 1;
            int de = it.max_depth, tail_levels = 0;
            while (de > d0[i] && wd[de] <= e->hybrid_width) tail_levels += wd[de--] > 0;
            if (tail_levels >= 2) d1[i] = de + 1;
        }
    }
    // node lists per (level) of the per-level launches
    std::vector<std::vector<std::pair<int, int>>> by_h(maxh + 1), by_d(maxd + 1);
    for (int i = 0; i < nitems; ++i) {
        const ItemHost& it = e->items[i];
        for (int v = 0; v < it.N; ++v) {
            if (it.nodes[v].nchild == 0) continue;
            if (it.height[v] < h0[i]) by_h[it.height[v]].push_back({i, v});
            if (it.depth[v] > d0[i] && it.depth[v] < d1[i]) by_d[it.depth[v]].push_back({i, v});
        }
    }
    // algorithmic bytes of one job (the accounting of mcp_get_bytes, by job):
    // post-order, node v reads its internal children's slabs (1 byte per
    // pattern for a leaf child) and writes its own unless it is the root;
    // pre-order, parent v reads its own pre-order slab unless it is the root,
    // and per child reads the partial and writes the pre-order slab (a leaf
    // child: its code)
    // (lds: the fused block keeps the slabs passed between its steps in LDS:
    // the post-order pass writes the partials and reads only tip codes, the
    // pre-order pass reads the partials and tip codes)
    auto job_bytes = [&](bool adjoint, int i, int v, bool lds = false) -> double {
        const ItemHost& it = e->items[i];
        const DatasetHost& d = e->ds[it.dataset];
        const double w = 8.0 * it.S * it.R * d.C;
        const NodeInfo& n = it.nodes[v];
        double b = (!n.is_root && !(adjoint && lds)) ? w : 0.0;
        for (int q = it.child_ptr[v]; q < it.child_ptr[v + 1]; ++q) {
            const bool inner = it.nodes[it.children[q]].nchild > 0;
            b += !inner ? (double)d.C : (adjoint ? (lds ? w : 2.0 * w) : (lds ? 0.0 : w));
        }
        return b;
    };
    // LDS stages of the fused kernel, per item: its leaves' tip-code rows of
    // the block's tile (when they fit kLdsCodesMax and its tiles stay within
    // the padded pattern range), then, under the fused schedule (every step
    // of the item in the one launch), a slab stage when the whole stage fits
    // kLdsFusedMax: the post-order pass keeps copies of its frontier's
    // partials there (slots NodeInfo::fl), the pre-order pass its pre-order
    // slabs (slots NodeInfo::gslot).
    std::vector<int> lds_rows(nitems, 0), lds_g(nitems, -1);
    std::vector<size_t> lds_end(nitems, 0);
    for (int i = 0; i < nitems; ++i) {
        ItemHost& it = e->items[i];
        const DatasetHost& d = e->ds[it.dataset];
        const size_t tcols = 64 * (size_t)std::max(1, e->cpt(it.S));
        const size_t cbytes = (size_t)d.nleaves * tcols;
        if (d.nleaves > 0 && cbytes <= kLdsCodesMax && (long long)it.ntiles * tcols <= d.Cp) {
            lds_rows[i] = d.nleaves;
            lds_end[i] = cbytes;
        }
        it.lds_stage = false;
        if (e->schedule == MCP_SCHEDULE_FUSED && e->lds_slabs) {
            const size_t goff = round_up(lds_end[i], 16);
            const size_t gbytes =
                (size_t)std::max(it.g_slots[0], it.fl_slots) * it.S * tcols * sizeof(double);
            if (goff + gbytes <= kLdsFusedMax) {
                lds_g[i] = (int)goff;
                lds_end[i] = goff + gbytes;
                it.lds_stage = true;
                // mcp_get_bytes: no pre-order slab traffic, and no re-read of
                // a partial by its mother's post-order task
                const double w = 8.0 * it.S * it.R * d.C;
                for (int v = 0; v < it.N; ++v)
                    if (it.nodes[v].nchild > 0 && !it.nodes[v].is_root) {
                        e->bytes_adj -= 2.0 * w;
                        e->bytes_fwd -= w;
                    }
            }
        }
    }
    auto add_launch = [&](bool adjoint, int S, const std::vector<std::pair<int, int>>& nodes,
                          std::vector<Launch>* target = nullptr) -> int {
        std::vector<LevelJob>& js = adjoint ? ja : jf;
        Launch L;
        L.adjoint = adjoint;
        L.S = S;
        L.cpt = e->cpt(S);
        L.bytes = 0.0;
        L.job_begin = (int)js.size();
        L.taskoff_begin = (int)e->taskoff.size();
        long long tasks = 0;
        e->taskoff.push_back(0);
        for (auto [i, v] : nodes) {
            const ItemHost& it = e->items[i];
            if (it.S != S) continue;
            js.push_back(make_job(i, v));
            L.bytes += job_bytes(adjoint, i, v);
            tasks += (long long)it.R * it.ntiles;
            e->taskoff.push_back((int32_t)tasks);
        }
        L.njobs = (int)js.size() - L.job_begin;
        L.ntasks = (int)tasks;
        if (L.njobs == 0) {
            e->taskoff.pop_back();
            return MCP_OK;
        }
        if (tasks > 0x7fffffffLL) return fail(MCP_ERR_UNSUPPORTED, "too many tasks in one level");
        // persistent grid: chunks of consecutive tasks, no empty block
        const int maxgrid = e->num_cu * MCP_GRID_PER_CU;
        int grid = (int)std::min<long long>(tasks, maxgrid);
        L.chunk = (int)((tasks + grid - 1) / grid);
        L.grid = (int)((tasks + L.chunk - 1) / L.chunk);
        L.blk_begin = (int)e->blkjob.size();
        const int32_t* off = e->taskoff.data() + L.taskoff_begin;
        int j = 0;
        for (int b = 0; b < L.grid; ++b) {
            const long long t0 = (long long)b * L.chunk;
            while (off[j + 1] <= t0) ++j;
            e->blkjob.push_back(j);
        }
        (target ? *target : (adjoint ? e->adj : e->fwd)).push_back(L);
        return MCP_OK;
    };
    for (int h = 1; h <= maxh; ++h)
        for (int S : svals)
            if ((st = add_launch(false, S, by_h[h]))) return st;
    for (int dd = 0; dd <= maxd; ++dd)
        for (int S : svals)
            if ((st = add_launch(true, S, by_d[dd]))) return st;
    // model-gradient pass: every pre-order level, one k_level_mg launch per
    // (depth, state count); under the fused schedule per pre-order Hu step
    // (the order its pre-order slabs are laid out for)
    e->adj_mg.clear();
    if (e->mg_ok) {
        std::vector<std::vector<std::pair<int, int>>> all_d;
        if (e->schedule == MCP_SCHEDULE_FUSED) {
            size_t maxs = 0;
            for (int i = 0; i < nitems; ++i) maxs = std::max(maxs, hu[i].size());
            all_d.resize(maxs);
            for (int i = 0; i < nitems; ++i) {
                const int ns = (int)hu[i].size();
                for (int k = 0; k < ns; ++k)
                    for (int v : hu[i][k]) all_d[ns - 1 - k].push_back({i, v});
            }
        } else {
            all_d.resize(maxd + 1);
            for (int i = 0; i < nitems; ++i) {
                const ItemHost& it = e->items[i];
                for (int v = 0; v < it.N; ++v)
                    if (it.nodes[v].nchild > 0) all_d[it.depth[v]].push_back({i, v});
            }
        }
        for (size_t dd = 0; dd < all_d.size(); ++dd)
            for (int S : svals)
                if ((st = add_launch(true, S, all_d[dd], &e->adj_mg))) return st;
    }
    if (e->schedule != MCP_SCHEDULE_LEVELS) {
        // fused: jobs grouped by item, then level; per-item level offsets
        std::vector<int32_t> lf, la;
        std::vector<ItemSched> sched;
        // per state count: the items' records and the end of their LDS stages
        std::vector<std::vector<std::pair<ItemSched, size_t>>> per_s(svals.size());
        std::vector<std::pair<double, double>> fbytes(svals.size(), {0.0, 0.0});
        for (int i = 0; i < nitems; ++i) {
            const ItemHost& it = e->items[i];
            const DatasetHost& d = e->ds[it.dataset];
            ItemSched I{};
            I.R = it.R;
            I.ntiles = it.ntiles;
            I.codes = e->dcodes.as<uint8_t>() + d.codes_off;
            I.Cp = d.Cp;
            I.lds_leaves = lds_rows[i];
            I.lds_g = lds_g[i];
            std::vector<std::vector<int>> hv(it.max_height + 1), dv(it.max_depth + 1);
            for (int v = 0; v < it.N; ++v) {
                if (it.nodes[v].nchild == 0) continue;
                hv[it.height[v]].push_back(v);
                dv[it.depth[v]].push_back(v);
            }
            const size_t ks = std::find(svals.begin(), svals.end(), it.S) - svals.begin();
            if (e->schedule == MCP_SCHEDULE_FUSED) {
                // the whole walk in one launch: Hu steps of at most `waves`
                // nodes, backwards for the pre-order pass
                const std::vector<std::vector<int>>& steps = hu[i];
                I.lf0 = (int)lf.size();
                for (const auto& step : steps) {
                    lf.push_back((int32_t)jf.size());
                    for (int v : step) {
                        jf.push_back(make_job(i, v));
                        fbytes[ks].first += job_bytes(false, i, v, lds_g[i] >= 0);
                    }
                }
                I.lf1 = (int)lf.size();
                I.la0 = (int)la.size();
                for (auto st2 = steps.rbegin(); st2 != steps.rend(); ++st2) {
                    la.push_back((int32_t)ja.size());
                    for (int v : *st2) {
                        ja.push_back(make_job(i, v));
                        fbytes[ks].second += job_bytes(true, i, v, lds_g[i] >= 0);
                    }
                }
                I.la1 = (int)la.size();
                fused_first[i] = steps.empty() ? -1 : steps[0][0];
            } else {
                I.lf0 = (int)lf.size();
                for (int h = h0[i]; h <= it.max_height; ++h) {
                    if (hv[h].empty()) continue;
                    lf.push_back((int32_t)jf.size());
                    for (int v : hv[h]) {
                        jf.push_back(make_job(i, v));
                        fbytes[ks].first += job_bytes(false, i, v);
                    }
                }
                I.lf1 = (int)lf.size();
                I.la0 = (int)la.size();
                for (int dd = 0; dd <= d0[i]; ++dd) {
                    if (dv[dd].empty()) continue;
                    la.push_back((int32_t)ja.size());
                    for (int v : dv[dd]) {
                        ja.push_back(make_job(i, v));
                        fbytes[ks].second += job_bytes(true, i, v);
                    }
                }
                I.la1 = (int)la.size();
            }
            const size_t k = std::find(svals.begin(), svals.end(), it.S) - svals.begin();
            per_s[k].push_back({I, lds_end[i]});
        }
        // tails: the deep narrow pre-order levels of each item that has them
        std::vector<std::vector<std::pair<ItemSched, size_t>>> tail_s(svals.size());
        std::vector<double> tbytes(svals.size(), 0.0);
        for (int i = 0; i < nitems; ++i) {
            const ItemHost& it = e->items[i];
            if (d1[i] > it.max_depth) continue;
            std::vector<std::vector<int>> dv(it.max_depth + 1);
            for (int v = 0; v < it.N; ++v)
                if (it.nodes[v].nchild > 0) dv[it.depth[v]].push_back(v);
            const size_t k = std::find(svals.begin(), svals.end(), it.S) - svals.begin();
            const DatasetHost& d = e->ds[it.dataset];
            ItemSched I{};
            I.R = it.R;
            I.ntiles = it.ntiles;
            I.codes = e->dcodes.as<uint8_t>() + d.codes_off;
            I.Cp = d.Cp;
            I.lds_leaves = lds_rows[i];
            I.lds_g = -1;
            I.lf0 = I.lf1 = 0;
            I.la0 = (int)la.size();
            for (int dd = d1[i]; dd <= it.max_depth; ++dd) {
                if (dv[dd].empty()) continue;
                la.push_back((int32_t)ja.size());
                for (int v : dv[dd]) {
                    ja.push_back(make_job(i, v));
                    tbytes[k] += job_bytes(true, i, v);
                }
            }
            I.la1 = (int)la.size();
            tail_s[k].push_back({I, (size_t)lds_rows[i] * 64 * std::max(1, e->cpt(it.S))});
        }
        lf.push_back((int32_t)jf.size());
        la.push_back((int32_t)ja.size());
        auto add_fused = [&](const std::vector<std::pair<ItemSched, size_t>>& items_k, int S,
                             double bf, double ba, bool tail) -> int {
            mcp_engine::FusedLaunch F{};
            F.S = S;
            F.cpt = e->cpt(S);
            F.bytes_f = bf;
            F.bytes_a = ba;
            F.tail = tail;
            F.sched_begin = (int)sched.size();
            F.lds_bytes = 0;
            long long blocks = 0;
            for (auto [I, lend] : items_k) {
                I.block_off = (int32_t)blocks;
                blocks += (long long)I.R * I.ntiles;
                F.lds_bytes = std::max(F.lds_bytes, lend);
                sched.push_back(I);
            }
            if (blocks > 0x7fffffffLL) return fail(MCP_ERR_UNSUPPORTED, "too many tiles");
            F.nitems = (int)items_k.size();
            F.nblocks = (int)blocks;
            if (F.nitems) e->fused.push_back(F);
            return MCP_OK;
        };
        for (size_t k = 0; k < svals.size(); ++k)
            if ((st = add_fused(per_s[k], svals[k], fbytes[k].first, fbytes[k].second, false)))
                return st;
        for (size_t k = 0; k < svals.size(); ++k)
            if ((st = add_fused(tail_s[k], svals[k], 0.0, tbytes[k], true))) return st;
        if ((st = upload(e, e->dsched, sched.data(), sched.size() * sizeof(ItemSched)))) return st;
        if ((st = upload(e, e->dlvl_f, lf.data(), lf.size() * sizeof(int32_t)))) return st;
        if ((st = upload(e, e->dlvl_a, la.data(), la.size() * sizeof(int32_t)))) return st;
    }
    if ((st = upload(e, e->djobs_f, jf.data(), jf.size() * sizeof(LevelJob)))) return st;
    if ((st = upload(e, e->djobs_a, ja.data(), ja.size() * sizeof(LevelJob)))) return st;
    if (!more.empty() && (st = upload(e, e->dmore, more.data(), more.size() * sizeof(ChildRef))))
        return st;
    if ((st = upload(e, e->dtaskoff, e->taskoff.data(), e->taskoff.size() * sizeof(int32_t))))
        return st;
    if ((st = upload(e, e->dblkjob, e->blkjob.data(), e->blkjob.size() * sizeof(int32_t))))
        return st;

    // reductions: ll of every item first, then every branch gradient; sums
    // longer than kRedChunk are cut into fixed chunks (stage 1 into a
    // temporary, stage 2 over the chunk sums) so no block sums millions
    e->red.clear();
    e->red2.clear();
    // 4096: a stage-1 block then walks its slice in two rounds of 8
    // independent loads per thread (32768 took 16 rounds: the straggler of
    // the launch)
    const int32_t kRedChunk = 1 << 12;
    int32_t ntmp = 0;
    auto add_sum = [&](int32_t dst, int32_t start, int32_t len) {
        if (len <= kRedChunk) {
            e->red.push_back(RedOp{dst, start, len, 0});
            return;
        }
        const int32_t first = ntmp;
        for (int32_t a = 0; a < len; a += kRedChunk)
            e->red.push_back(RedOp{ntmp++, start + a, std::min(kRedChunk, len - a), 1});
        e->red2.push_back(RedOp{dst, first, ntmp - first, 0});
    };
    for (int i = 0; i < nitems; ++i) {
        const ItemHost& it = e->items[i];
        int v0 = -1;  // the fused launch's first post-order job, when it runs them all
        if (e->schedule == MCP_SCHEDULE_FUSED) {
            v0 = fused_first[i];
        } else if (e->schedule == MCP_SCHEDULE_HYBRID && h0[i] == 1) {
            for (int v = 0; v < it.N && v0 < 0; ++v)
                if (it.nodes[v].nchild > 0 && it.height[v] == 1) v0 = v;
        }
        if (v0 >= 0) {
            // the fused tasks store 0 into their log-scaler slots and each block
            // puts its whole sum into the first job's slot: only those count
            add_sum(it.out_off, it.parts_off + it.nodes[v0].lpart, it.R * it.ntiles);
        } else {
            add_sum(it.out_off, it.parts_off, it.nint * it.R * it.ntiles);
        }
    }
    e->n_ll_ops = (int)e->red.size();
    e->n_ll_ops2 = (int)e->red2.size();
    for (auto& it : e->items)
        for (int v = 0; v < it.N - 1; ++v)
            add_sum(it.out_off + 1 + v, it.parts_off + (it.nint + v) * it.R * it.ntiles,
                    it.R * it.ntiles);
    if (!e->red2.empty()) {
        if ((st = upload(e, e->dred2, e->red2.data(), e->red2.size() * sizeof(RedOp)))) return st;
        if ((st = e->dredtmp.ensure((size_t)ntmp * sizeof(double)))) return st;
    }
    e->red_vr.clear();
    e->vr_off.assign(e->items.size(), 0);
    e->total_vr = 0;
    for (size_t i = 0; i < e->items.size(); ++i) {
        const ItemHost& it = e->items[i];
        e->vr_off[i] = e->total_vr;
        for (int v = 0; v < it.N - 1; ++v)
            for (int r = 0; r < it.R; ++r)
                e->red_vr.push_back(RedOp{(int32_t)(e->total_vr + (size_t)v * it.R + r),
                                          it.parts_off + (it.nint + v) * it.R * it.ntiles +
                                              r * it.ntiles,
                                          it.ntiles, 1});
        e->total_vr += (size_t)(it.N - 1) * it.R;
    }
    if ((st = upload(e, e->dred_vr, e->red_vr.data(), e->red_vr.size() * sizeof(RedOp)))) return st;
    if ((st = e->dout_vr.ensure(std::max<size_t>(e->total_vr, 1) * sizeof(double)))) return st;
    if ((st = e->hout_vr.ensure(std::max<size_t>(e->total_vr, 1) * sizeof(double)))) return st;
    e->red_mg.clear();
    e->mg_out_off.assign(e->items.size(), 0);
    e->total_mg = 0;
    if (e->mg_ok) {
        for (size_t i = 0; i < e->items.size(); ++i) {
            const ItemHost& it = e->items[i];
            const int SS = it.S * it.S;
            e->mg_out_off[i] = e->total_mg;
            // per (branch, r, e): the tiles of mparts[mg_off + ((v R + r) SS + e) ntiles ..]
            for (long long q = 0; q < (long long)(it.N - 1) * it.R * SS; ++q)
                e->red_mg.push_back(RedOp{(int32_t)(e->total_mg + q),
                                          (int32_t)(it.mg_off + q * it.ntiles), it.ntiles, 1});
            const size_t root = e->total_mg + (size_t)(it.N - 1) * it.R * SS;
            const long long rbase = it.mg_off + (long long)(it.N - 1) * it.R * SS * it.ntiles;
            // per state: the (r, tile) block of the root term
            for (int s2 = 0; s2 < it.S; ++s2)
                e->red_mg.push_back(RedOp{(int32_t)(root + s2),
                                          (int32_t)(rbase + (long long)s2 * it.R * it.ntiles),
                                          it.R * it.ntiles, 1});
            e->total_mg += (size_t)(it.N - 1) * it.R * SS + it.S;
        }
        if ((st = upload(e, e->dred_mg, e->red_mg.data(), e->red_mg.size() * sizeof(RedOp))))
            return st;
        if ((st = e->dout_mg.ensure(std::max<size_t>(e->total_mg, 1) * sizeof(double)))) return st;
        if ((st = e->hout_mg.ensure(std::max<size_t>(e->total_mg, 1) * sizeof(double)))) return st;
    }
    if ((st = upload(e, e->dred, e->red.data(), e->red.size() * sizeof(RedOp)))) return st;
    return MCP_OK;
}

// Offsets of the model-gradient partial sums (per item: [branch][r][S*S][tile]
// then the root's [S][r][tile]); the pass exists only when every item has
// model-gradient kernels for its tile width and the offsets fit 32 bits.
static void assign_mg_offsets(mcp_engine* e) {
    long long tot = 0;
    bool ok = true;
    for (auto& it : e->items) {
        if (!mg_available(it.S, e->cpt(it.S))) ok = false;
        it.mg_off = tot;
        tot += ((long long)(it.N - 1) * it.R * it.S * it.S + (long long)it.S * it.R) * it.ntiles;
    }
    if (tot > 0x7fffffffLL) ok = false;
    e->mg_ok = ok;
    e->total_mparts = ok ? tot : 0;
    if (!ok)
        for (auto& it : e->items) it.mg_off = -1;
}

static int engine_create(mcp_engine** out, const mcp_batch_desc* d) {
    if (!out || !d) return fail(MCP_ERR_INVALID, "null argument");
    *out = nullptr;
    if (d->n_datasets < 1 || d->n_items < 1 || !d->datasets || !d->items)
        return fail(MCP_ERR_INVALID, "batch needs at least one dataset and one item");
#ifndef MCP_FUSED_WAVES
#define MCP_FUSED_WAVES 4
#endif
    const int waves = d->waves_per_block ? d->waves_per_block : MCP_FUSED_WAVES;
    if (waves < 1 || waves > 8) return fail(MCP_ERR_INVALID, "waves_per_block must be 1..8");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (d->device < 0 || d->device >= ndev)
        return fail(MCP_ERR_INVALID, "device ordinal out of range");
    DeviceGuard guard(d->device);
    std::unique_ptr<mcp_engine> e(new (std::nothrow) mcp_engine());
    if (!e) return fail(MCP_ERR_NOMEM, "host allocation failed");
    e->device = d->device;
    e->waves = waves;
    if (const char* v = getenv("MCP_FUSED_LDS_SLABS")) e->lds_slabs = atoi(v) != 0;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, d->device));
    e->num_cu = prop.multiProcessorCount;
    HIP_TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    int st;
    e->ds.resize(d->n_datasets);
    for (int k = 0; k < d->n_datasets; ++k)
        if ((st = build_dataset(d->datasets[k], e->ds[k]))) return st;
    e->items.resize(d->n_items);
    for (int i = 0; i < d->n_items; ++i)
        if ((st = item_from_desc(d->items[i], e->ds, e->items[i]))) return st;

    // fixed layout
    size_t codes = 0, tiptab = 0, weights = 0;
    for (auto& D : e->ds) {
        D.codes_off = codes;
        D.tiptab_off = tiptab;
        D.weights_off = weights;
        codes += D.codes.size();
        tiptab += D.tiptab.size();
        weights += D.weights.size();
    }
    size_t F = 0, P = 0, PT = 0, model = 0, blv = 0;
    int parts = 0, outs = 0;
    e->total_patterns = 0;
    for (auto& it : e->items) {
        const DatasetHost& D = e->ds[it.dataset];
        it.F_off = F;
        F += (size_t)(it.nint - 1) * it.R * D.Cp * it.S;
        it.P_off = P;
        P += (size_t)(it.N - 1) * it.R * it.S * it.S;
        it.PT_off = PT;
        PT += (size_t)D.nleaves * it.R * D.K * it.S;
        it.model_off = model;
        model += 2 * (size_t)it.S * it.S + 2 * (size_t)it.S + it.R;
        it.blv_off = blv;
        blv += it.N - 1;
        const long long nparts = (long long)(it.nint + it.N - 1) * it.R * it.ntiles;
        if ((long long)parts + nparts > 0x7fffffffLL)
            return fail(MCP_ERR_UNSUPPORTED, "batch too large for 32-bit partial-sum offsets");
        it.parts_off = parts;
        parts += (int)nparts;
        it.out_off = outs;
        outs += it.N;
        e->total_patterns += D.C;
    }
    e->total_blv = blv;
    e->total_out = outs;
    if ((st = e->dF.ensure(std::max<size_t>(F, 1) * sizeof(double)))) return st;
    if ((st = e->dP.ensure(P * sizeof(double)))) return st;
    if ((st = e->ddP.ensure(P * sizeof(double)))) return st;
    if ((st = e->dPT.ensure(std::max<size_t>(PT, 1) * sizeof(double)))) return st;
    if ((st = e->ddPT.ensure(std::max<size_t>(PT, 1) * sizeof(double)))) return st;
    if ((st = e->dmodel.ensure(model * sizeof(double)))) return st;
    if ((st = e->dblv.ensure(blv * sizeof(double)))) return st;
    if ((st = e->dparts.ensure((size_t)parts * sizeof(double)))) return st;
    if ((st = e->dout.ensure((size_t)outs * sizeof(double)))) return st;
    if ((st = e->hblv.ensure(blv * sizeof(double)))) return st;
    if ((st = e->hout.ensure((size_t)outs * sizeof(double)))) return st;
    HIP_TRY(hipMemset(e->dparts.p, 0, (size_t)parts * sizeof(double)));
    {
        std::vector<uint8_t> hc(codes);
        std::vector<double> ht(tiptab), hw(weights);
        for (auto& D : e->ds) {
            std::copy(D.codes.begin(), D.codes.end(), hc.begin() + D.codes_off);
            std::copy(D.tiptab.begin(), D.tiptab.end(), ht.begin() + D.tiptab_off);
            std::copy(D.weights.begin(), D.weights.end(), hw.begin() + D.weights_off);
        }
        if ((st = upload(e.get(), e->dcodes, hc.data(), hc.size()))) return st;
        if ((st = upload(e.get(), e->dtiptab, ht.data(), ht.size() * sizeof(double)))) return st;
        if ((st = upload(e.get(), e->dweights, hw.data(), hw.size() * sizeof(double)))) return st;
    }
    assign_mg_offsets(e.get());
    if ((st = rebuild_schedule(e.get()))) return st;
    *out = e.release();
    return MCP_OK;
}

// Tile widths changed: every item's tile count, the partial-sum slots and
// the schedule follow (device slabs keep their layout: the pattern axis is
// padded to the widest candidate).
static int relayout_tiles(mcp_engine* e) {
    long long parts = 0;
    for (auto& it : e->items) {
        const DatasetHost& D = e->ds[it.dataset];
        it.ntiles = D.Cp / tile_cols(e->cpt(it.S));
        assign_partial_slots(it);
        const long long nparts = (long long)(it.nint + it.N - 1) * it.R * it.ntiles;
        if (parts + nparts > 0x7fffffffLL)
            return fail(MCP_ERR_UNSUPPORTED, "batch too large for 32-bit partial-sum offsets");
        it.parts_off = (int)parts;
        parts += nparts;
    }
    int st;
    if ((st = e->dparts.ensure((size_t)std::max<long long>(parts, 1) * sizeof(double)))) return st;
    HIP_TRY(hipMemset(e->dparts.p, 0, (size_t)std::max<long long>(parts, 1) * sizeof(double)));
    assign_mg_offsets(e);
    return rebuild_schedule(e);
}

static int engine_set_model(mcp_engine* e, int item, const double* U, const double* D,
                            const double* Uinv, double mu, const double* rates, const double* pi) {
    if (!e) return fail(MCP_ERR_INVALID, "null handle");
    if (item < 0 || item >= (int)e->items.size()) return fail(MCP_ERR_INVALID, "item out of range");
    if (!U || !D || !Uinv || !rates || !pi) return fail(MCP_ERR_INVALID, "null model array");
    ItemHost& it = e->items[item];
    const int S = it.S, R = it.R;
    std::vector<double> buf;
    buf.insert(buf.end(), U, U + S * S);
    buf.insert(buf.end(), D, D + S);
    buf.insert(buf.end(), Uinv, Uinv + S * S);
    buf.insert(buf.end(), rates, rates + R);
    buf.insert(buf.end(), pi, pi + S);
    for (double x : buf)
        if (!std::isfinite(x)) return fail(MCP_ERR_COMPLEX, "model arrays must be finite");
    if (!std::isfinite(mu)) return fail(MCP_ERR_COMPLEX, "mu must be finite");
    DeviceGuard guard(e->device);
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpy(e->dmodel.as<double>() + it.model_off, buf.data(), buf.size() * sizeof(double),
                      hipMemcpyHostToDevice));
    const bool mu_changed = !it.model_set || it.mu != mu;
    it.U.assign(U, U + S * S);
    it.D.assign(D, D + S);
    it.Uinv.assign(Uinv, Uinv + S * S);
    it.rates.assign(rates, rates + R);
    it.pi.assign(pi, pi + S);
    it.mu = mu;
    it.model_set = true;
    if (mu_changed) {
        // mu lives in the device ModelDev record
        ModelDev md;
        HIP_TRY(hipMemcpy(&md, e->dmodels.as<ModelDev>() + item, sizeof(ModelDev),
                          hipMemcpyDeviceToHost));
        md.mu = mu;
        HIP_TRY(hipMemcpy(e->dmodels.as<ModelDev>() + item, &md, sizeof(ModelDev),
                          hipMemcpyHostToDevice));
    }
    return MCP_OK;
}

static int engine_set_topology(mcp_engine* e, int item, const int32_t* parent) {
    if (!e || !parent) return fail(MCP_ERR_INVALID, "null argument");
    if (item < 0 || item >= (int)e->items.size()) return fail(MCP_ERR_INVALID, "item out of range");
    DeviceGuard guard(e->device);
    HIP_TRY(hipStreamSynchronize(e->stream));
    ItemHost tmp = e->items[item];
    int st = build_topology(tmp, e->ds[tmp.dataset], parent);
    if (st) return st;
    if (tmp.nint != e->items[item].nint)
        return fail(MCP_ERR_DIMENSION, "new topology changes the number of internal nodes");
    e->items[item] = std::move(tmp);
    st = rebuild_schedule(e);
    if (st) return st;
    // the ModelDev records were rebuilt from host copies; mu included
    return MCP_OK;
}

// Enqueue one evaluation on the engine's stream (no synchronisation).  In
// timing mode (eager launches only) every kernel launch is bracketed by a
// pair of HIP events on the same stream.
static int enqueue_eval(mcp_engine* e, bool grad, bool by_rate = false) {
    hipStream_t st = e->stream;
    const bool tm = e->timing && !e->use_graph;
    e->nk = 0;
    double launch_bytes = 0.0;  // algorithmic bytes of the next timed launch
    auto before = [&](int phase) -> int {
        if (!tm) return MCP_OK;
        const size_t k = (size_t)e->nk;
        while (e->kev.size() < 2 * k + 2) {
            hipEvent_t ev;
            HIP_TRY(hipEventCreate(&ev));
            e->kev.push_back(ev);
        }
        if (e->kphase.size() < k + 1) e->kphase.resize(k + 1);
        if (e->kbytes.size() < k + 1) e->kbytes.resize(k + 1);
        e->kphase[k] = phase;
        e->kbytes[k] = launch_bytes;
        launch_bytes = 0.0;
        HIP_TRY(hipEventRecord(e->kev[2 * k], st));
        return MCP_OK;
    };
    auto after = [&]() -> int {
        if (!tm) return MCP_OK;
        HIP_TRY(hipEventRecord(e->kev[2 * (size_t)e->nk + 1], st));
        e->nk++;
        return MCP_OK;
    };
    int rc;
#define TIMED(phase, call)                 \
    if ((rc = before(phase))) return rc;   \
    HIP_TRY(call);                         \
    if ((rc = after())) return rc;
    if (tm) HIP_TRY(hipEventRecord(e->ev_begin, st));
    HIP_TRY(hipMemcpyAsync(e->dblv.p, e->hblv.p, e->total_blv * sizeof(double),
                           hipMemcpyHostToDevice, st));
    TIMED(1, launch_build_P_tips(e->dmodels.as<ModelDev>(), e->delemP.as<int32_t>(),
                                 e->delemT.as<int32_t>(), (int)e->items.size(), e->nelemP,
                                 e->nelemT, st));
    for (const Launch& L : e->fwd) {
        launch_bytes = L.bytes;
        TIMED(2, launch_level(false, L.S, L.cpt, e->djobs_f.as<LevelJob>() + L.job_begin,
                              e->dtaskoff.as<int32_t>() + L.taskoff_begin,
                              e->dblkjob.as<int32_t>() + L.blk_begin, L.ntasks, L.chunk, L.grid,
                              e->dparts.as<double>(), st));
    }
    // the fused levels sit between the two passes (top of the post-order
    // pass, start of the pre-order pass)
    for (const auto& F : e->fused) {
        if (F.tail) continue;
        launch_bytes = F.bytes_f + (grad ? F.bytes_a : 0.0);
        TIMED(grad ? 3 : 2,
              launch_fused(grad, F.S, F.cpt, e->dsched.as<ItemSched>() + F.sched_begin, F.nitems,
                           F.nblocks, e->waves, F.lds_bytes, e->djobs_f.as<LevelJob>(), e->dlvl_f.as<int32_t>(),
                           e->djobs_a.as<LevelJob>(), e->dlvl_a.as<int32_t>(),
                           e->dparts.as<double>(), st));
    }
    if (grad) {
        for (const Launch& L : e->adj) {
            launch_bytes = L.bytes;
            TIMED(3, launch_level(true, L.S, L.cpt, e->djobs_a.as<LevelJob>() + L.job_begin,
                                  e->dtaskoff.as<int32_t>() + L.taskoff_begin,
                                  e->dblkjob.as<int32_t>() + L.blk_begin, L.ntasks, L.chunk,
                                  L.grid, e->dparts.as<double>(), st));
        }
        for (const auto& F : e->fused) {
            if (!F.tail) continue;
            launch_bytes = F.bytes_a;
            TIMED(3, launch_fused(true, F.S, F.cpt, e->dsched.as<ItemSched>() + F.sched_begin,
                                  F.nitems, F.nblocks, e->waves, F.lds_bytes, e->djobs_f.as<LevelJob>(),
                                  e->dlvl_f.as<int32_t>(), e->djobs_a.as<LevelJob>(),
                                  e->dlvl_a.as<int32_t>(), e->dparts.as<double>(), st));
        }
    }
    const int nops = (grad && !by_rate) ? (int)e->red.size() : e->n_ll_ops;
    const int nops2 = (grad && !by_rate) ? (int)e->red2.size() : e->n_ll_ops2;
    TIMED(4, launch_reduce(e->dred.as<RedOp>(), nops, e->dparts.as<double>(),
                           e->dout.as<double>(), e->dredtmp.as<double>(), st));
    if (nops2 > 0) {
        TIMED(4, launch_reduce(e->dred2.as<RedOp>(), nops2, e->dredtmp.as<double>(),
                               e->dout.as<double>(), e->dredtmp.as<double>(), st));
    }
    if (grad && by_rate) {
        TIMED(4, launch_reduce(e->dred_vr.as<RedOp>(), (int)e->red_vr.size(),
                               e->dparts.as<double>(), e->dout.as<double>(),
                               e->dout_vr.as<double>(), st));
        HIP_TRY(hipMemcpyAsync(e->hout_vr.p, e->dout_vr.p, e->total_vr * sizeof(double),
                               hipMemcpyDeviceToHost, st));
    }
#undef TIMED
    HIP_TRY(hipMemcpyAsync(e->hout.p, e->dout.p, e->total_out * sizeof(double),
                           hipMemcpyDeviceToHost, st));
    if (tm) HIP_TRY(hipEventRecord(e->ev_end, st));
    return MCP_OK;
}

static int engine_eval(mcp_engine* e, const double* blv, double* ll, double* grad, bool want_grad) {
    if (!e || !blv || !ll) return fail(MCP_ERR_INVALID, "null argument");
    for (auto& it : e->items)
        if (!it.model_set) return fail(MCP_ERR_STATE, "substitution model not set");
    DeviceGuard guard(e->device);
    std::memcpy(e->hblv.p, blv, e->total_blv * sizeof(double));
    e->evaluated = true;
    const int gi = want_grad ? 1 : 0;
    if (e->use_graph) {
        if (!e->graph_exec[gi]) {
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            int st = enqueue_eval(e, want_grad);
            hipGraph_t cap = nullptr;
            hipError_t ce = hipStreamEndCapture(e->stream, &cap);
            if (st) {
                if (cap) (void)hipGraphDestroy(cap);
                return st;
            }
            if (ce != hipSuccess) return fail(MCP_ERR_HIP, "graph capture failed");
            g = cap;
            hipError_t ie = hipGraphInstantiate(&e->graph_exec[gi], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                e->graph_exec[gi] = nullptr;
                return fail(MCP_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
            }
        }
        HIP_TRY(hipGraphLaunch(e->graph_exec[gi], e->stream));
    } else {
        int st = enqueue_eval(e, want_grad);
        if (st) return st;
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (e->timing && !e->use_graph) {
        float ms;
        HIP_TRY(hipEventElapsedTime(&ms, e->ev_begin, e->ev_end));
        e->last_ms[0] = ms;
        for (int k = 1; k < 5; ++k) e->last_ms[k] = 0.0;
        e->kms.assign(e->nk, 0.0);
        for (int k = 0; k < e->nk; ++k) {
            HIP_TRY(hipEventElapsedTime(&ms, e->kev[2 * k], e->kev[2 * k + 1]));
            e->kms[k] = ms;
            e->last_ms[e->kphase[k]] += ms;
        }
    }
    const double* o = e->hout.as<double>();
    size_t goff = 0;
    for (size_t i = 0; i < e->items.size(); ++i) {
        const ItemHost& it = e->items[i];
        ll[i] = o[it.out_off];
        if (want_grad && grad) std::memcpy(grad + goff, o + it.out_off + 1, (it.N - 1) * sizeof(double));
        goff += it.N - 1;
    }
    return MCP_OK;
}

// logL and the per-(branch, category) components of the gradient (eager
// launches): grad_vr[item][v][r], summing over r gives the branch gradient.
static int engine_eval_by_rate(mcp_engine* e, const double* blv, double* ll, double* grad_vr) {
    if (!e || !blv || !ll || !grad_vr) return fail(MCP_ERR_INVALID, "null argument");
    for (auto& it : e->items)
        if (!it.model_set) return fail(MCP_ERR_STATE, "substitution model not set");
    DeviceGuard guard(e->device);
    std::memcpy(e->hblv.p, blv, e->total_blv * sizeof(double));
    e->evaluated = true;
    const bool tm = e->timing;
    e->timing = false;
    int st = enqueue_eval(e, true, true);
    e->timing = tm;
    if (st) return st;
    HIP_TRY(hipStreamSynchronize(e->stream));
    const double* o = e->hout.as<double>();
    for (size_t i = 0; i < e->items.size(); ++i) ll[i] = o[e->items[i].out_off];
    std::memcpy(grad_vr, e->hout_vr.p, e->total_vr * sizeof(double));
    return MCP_OK;
}

// logL, branch-length gradient and the model gradient (eager launches):
// P build, the post-order pass of the current schedule (its fused launches
// without their pre-order levels), every pre-order level with k_level_mg,
// the reductions.  dP[item] is [branch][r][S][S] (P[s][s1] layout, row =
// parent state), dpi[item] is [S] (root term only).
static int engine_eval_model(mcp_engine* e, const double* blv, double* ll, double* grad,
                             double* dP, double* dpi) {
    if (!e || !blv || !ll || !grad || !dP || !dpi) return fail(MCP_ERR_INVALID, "null argument");
    for (auto& it : e->items)
        if (!it.model_set) return fail(MCP_ERR_STATE, "substitution model not set");
    if (!e->mg_ok)
        return fail(MCP_ERR_UNSUPPORTED,
                    "model gradient needs S <= 8 with a built tile width for every item");
    DeviceGuard guard(e->device);
    int st;
    if ((st = e->dmparts.ensure((size_t)std::max<long long>(e->total_mparts, 1) * sizeof(double))))
        return st;
    std::memcpy(e->hblv.p, blv, e->total_blv * sizeof(double));
    e->evaluated = true;
    hipStream_t s = e->stream;
    HIP_TRY(hipMemcpyAsync(e->dblv.p, e->hblv.p, e->total_blv * sizeof(double),
                           hipMemcpyHostToDevice, s));
    HIP_TRY(launch_build_P_tips(e->dmodels.as<ModelDev>(), e->delemP.as<int32_t>(),
                                e->delemT.as<int32_t>(), (int)e->items.size(), e->nelemP,
                                e->nelemT, s));
    for (const Launch& L : e->fwd)
        HIP_TRY(launch_level(false, L.S, L.cpt, e->djobs_f.as<LevelJob>() + L.job_begin,
                             e->dtaskoff.as<int32_t>() + L.taskoff_begin,
                             e->dblkjob.as<int32_t>() + L.blk_begin, L.ntasks, L.chunk, L.grid,
                             e->dparts.as<double>(), s));
    for (const auto& F : e->fused)
        if (!F.tail)
            HIP_TRY(launch_fused(false, F.S, F.cpt, e->dsched.as<ItemSched>() + F.sched_begin,
                                 F.nitems, F.nblocks, e->waves, F.lds_bytes, e->djobs_f.as<LevelJob>(),
                                 e->dlvl_f.as<int32_t>(), e->djobs_a.as<LevelJob>(),
                                 e->dlvl_a.as<int32_t>(), e->dparts.as<double>(), s));
    for (const Launch& L : e->adj_mg)
        HIP_TRY(launch_level_mg(L.S, L.cpt, e->djobs_a.as<LevelJob>() + L.job_begin,
                                e->dtaskoff.as<int32_t>() + L.taskoff_begin,
                                e->dblkjob.as<int32_t>() + L.blk_begin, L.ntasks, L.chunk, L.grid,
                                e->dparts.as<double>(), e->dmparts.as<double>(), s));
    HIP_TRY(launch_reduce(e->dred.as<RedOp>(), (int)e->red.size(), e->dparts.as<double>(),
                          e->dout.as<double>(), e->dredtmp.as<double>(), s));
    if (!e->red2.empty())
        HIP_TRY(launch_reduce(e->dred2.as<RedOp>(), (int)e->red2.size(), e->dredtmp.as<double>(),
                              e->dout.as<double>(), e->dredtmp.as<double>(), s));
    HIP_TRY(launch_reduce(e->dred_mg.as<RedOp>(), (int)e->red_mg.size(), e->dmparts.as<double>(),
                          e->dout.as<double>(), e->dout_mg.as<double>(), s));
    HIP_TRY(hipMemcpyAsync(e->hout.p, e->dout.p, e->total_out * sizeof(double),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(e->hout_mg.p, e->dout_mg.p, e->total_mg * sizeof(double),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const double* o = e->hout.as<double>();
    const double* om = e->hout_mg.as<double>();
    size_t goff = 0, poff = 0, soff = 0;
    for (size_t i = 0; i < e->items.size(); ++i) {
        const ItemHost& it = e->items[i];
        ll[i] = o[it.out_off];
        std::memcpy(grad + goff, o + it.out_off + 1, (it.N - 1) * sizeof(double));
        goff += it.N - 1;
        const size_t nP = (size_t)(it.N - 1) * it.R * it.S * it.S;
        std::memcpy(dP + poff, om + e->mg_out_off[i], nP * sizeof(double));
        poff += nP;
        std::memcpy(dpi + soff, om + e->mg_out_off[i] + nP, it.S * sizeof(double));
        soff += it.S;
    }
    return MCP_OK;
}

// ---------------------------------------------------------------------------
// device leapfrog (NUTS / HMC step over branch lengths)
// ---------------------------------------------------------------------------

static int hmc_prepare(mcp_engine* e) {
    const int n = (int)e->items.size();
    if (e->hmc_items.size() != (size_t)n) {
        e->hmc_items.assign(n, HmcItem{});
        e->hmc_items_dirty = true;
    }
    if (!e->hmc_items_dirty) return MCP_OK;
    for (int i = 0; i < n; ++i) {
        HmcItem& h = e->hmc_items[i];
        h.n = e->items[i].N - 1;
        h.off = (int32_t)e->items[i].blv_off;
        h.out = e->items[i].out_off;
    }
    int st;
    if ((st = upload(e, e->dhmc_items, e->hmc_items.data(), n * sizeof(HmcItem)))) return st;
    const size_t T = e->total_blv;
    {
        std::vector<uint8_t> inl(std::max<size_t>(T, 1), 0);
        for (int i = 0; i < n; ++i) {
            const ItemHost& it = e->items[i];
            for (int v = 0; v < it.N - 1; ++v)
                inl[it.blv_off + v] = it.nodes[v].nchild > 0 ? 1 : 0;
        }
        if ((st = upload(e, e->dhmc_int, inl.data(), inl.size()))) return st;
    }
    if ((st = e->dhmc.ensure((3 * T + 2 * (size_t)n) * sizeof(double)))) return st;
    if ((st = e->dlp.ensure((size_t)n * sizeof(double)))) return st;
    if ((st = e->hhmc.ensure((3 * T + 2 * (size_t)n + n + e->total_out) * sizeof(double)))) return st;
    e->hmc_items_dirty = false;
    e->drop_graphs();
    return MCP_OK;
}

// H2D [x|r|g|eps,delta] -> k_lf_pre -> P, tips, levels -> reduce -> k_lf_post
// -> D2H [x|r|g], lp, (ll, grad) outputs
static int enqueue_leapfrog(mcp_engine* e) {
    hipStream_t st = e->stream;
    const size_t T = e->total_blv;
    const int n = (int)e->items.size();
    double* dx = e->dhmc.as<double>();
    double* dr = dx + T;
    double* dg = dr + T;
    double* ded = dg + T;
    double* h = e->hhmc.as<double>();
    HIP_TRY(hipMemcpyAsync(dx, h, (3 * T + 2 * (size_t)n) * sizeof(double), hipMemcpyHostToDevice,
                           st));
    HIP_TRY(launch_lf_pre(e->dhmc_items.as<HmcItem>(), n, dx, dr, dg, ded, e->dblv.as<double>(),
                          64 * (int)std::max<size_t>(1, T), st));
    HIP_TRY(launch_build_P_tips(e->dmodels.as<ModelDev>(), e->delemP.as<int32_t>(),
                                e->delemT.as<int32_t>(), n, e->nelemP, e->nelemT, st));
    for (const Launch& L : e->fwd)
        HIP_TRY(launch_level(false, L.S, L.cpt, e->djobs_f.as<LevelJob>() + L.job_begin,
                             e->dtaskoff.as<int32_t>() + L.taskoff_begin,
                             e->dblkjob.as<int32_t>() + L.blk_begin, L.ntasks, L.chunk, L.grid,
                             e->dparts.as<double>(), st));
    for (const auto& F : e->fused)
        if (!F.tail)
            HIP_TRY(launch_fused(true, F.S, F.cpt, e->dsched.as<ItemSched>() + F.sched_begin,
                                 F.nitems, F.nblocks, e->waves, F.lds_bytes, e->djobs_f.as<LevelJob>(),
                                 e->dlvl_f.as<int32_t>(), e->djobs_a.as<LevelJob>(),
                                 e->dlvl_a.as<int32_t>(), e->dparts.as<double>(), st));
    for (const Launch& L : e->adj)
        HIP_TRY(launch_level(true, L.S, L.cpt, e->djobs_a.as<LevelJob>() + L.job_begin,
                             e->dtaskoff.as<int32_t>() + L.taskoff_begin,
                             e->dblkjob.as<int32_t>() + L.blk_begin, L.ntasks, L.chunk, L.grid,
                             e->dparts.as<double>(), st));
    for (const auto& F : e->fused)
        if (F.tail)
            HIP_TRY(launch_fused(true, F.S, F.cpt, e->dsched.as<ItemSched>() + F.sched_begin,
                                 F.nitems, F.nblocks, e->waves, F.lds_bytes, e->djobs_f.as<LevelJob>(),
                                 e->dlvl_f.as<int32_t>(), e->djobs_a.as<LevelJob>(),
                                 e->dlvl_a.as<int32_t>(), e->dparts.as<double>(), st));
    HIP_TRY(launch_reduce(e->dred.as<RedOp>(), (int)e->red.size(), e->dparts.as<double>(),
                          e->dout.as<double>(), e->dredtmp.as<double>(), st));
    HIP_TRY(launch_reduce(e->dred2.as<RedOp>(), (int)e->red2.size(), e->dredtmp.as<double>(),
                          e->dout.as<double>(), e->dredtmp.as<double>(), st));
    HIP_TRY(launch_lf_post(e->dhmc_items.as<HmcItem>(), n, dx, dr, dg, ded, e->dout.as<double>(),
                           e->dlp.as<double>(), e->dhmc_int.as<uint8_t>(), st));
    HIP_TRY(hipMemcpyAsync(h, dx, 3 * T * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(h + 3 * T + 2 * (size_t)n, e->dlp.p, n * sizeof(double),
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(h + 3 * T + 3 * (size_t)n, e->dout.p, e->total_out * sizeof(double),
                           hipMemcpyDeviceToHost, st));
    return MCP_OK;
}

static int engine_leapfrog(mcp_engine* e, const double* eps_delta, double* x, double* r, double* g,
                           double* lp, double* ll) {
    if (!e || !eps_delta || !x || !r || !g || !lp) return fail(MCP_ERR_INVALID, "null argument");
    for (auto& it : e->items)
        if (!it.model_set) return fail(MCP_ERR_STATE, "substitution model not set");
    DeviceGuard guard(e->device);
    int st;
    if ((st = hmc_prepare(e))) return st;
    const size_t T = e->total_blv;
    const int n = (int)e->items.size();
    double* h = e->hhmc.as<double>();
    std::memcpy(h, x, T * sizeof(double));
    std::memcpy(h + T, r, T * sizeof(double));
    std::memcpy(h + 2 * T, g, T * sizeof(double));
    std::memcpy(h + 3 * T, eps_delta, 2 * (size_t)n * sizeof(double));
    if (e->use_graph) {
        if (!e->graph_exec[2]) {
            HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            st = enqueue_leapfrog(e);
            hipGraph_t cap = nullptr;
            hipError_t ce = hipStreamEndCapture(e->stream, &cap);
            if (st) {
                if (cap) (void)hipGraphDestroy(cap);
                return st;
            }
            if (ce != hipSuccess) return fail(MCP_ERR_HIP, "graph capture failed");
            hipError_t ie = hipGraphInstantiate(&e->graph_exec[2], cap, nullptr, nullptr, 0);
            (void)hipGraphDestroy(cap);
            if (ie != hipSuccess) {
                e->graph_exec[2] = nullptr;
                return fail(MCP_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
            }
        }
        HIP_TRY(hipGraphLaunch(e->graph_exec[2], e->stream));
    } else {
        if ((st = enqueue_leapfrog(e))) return st;
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    std::memcpy(x, h, T * sizeof(double));
    std::memcpy(r, h + T, T * sizeof(double));
    std::memcpy(g, h + 2 * T, T * sizeof(double));
    std::memcpy(lp, h + 3 * T + 2 * (size_t)n, n * sizeof(double));
    if (ll) {
        const double* o = h + 3 * T + 3 * (size_t)n;
        for (int i = 0; i < n; ++i) ll[i] = o[e->items[i].out_off];
    }
    return MCP_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

const char* mcp_last_error(void) { return g_last_error.c_str(); }

int mcp_plan_create(mcp_plan** out, const mcp_plan_desc* d) {
    if (!out || !d) return fail(MCP_ERR_INVALID, "null argument");
    if (!d->parent || !d->tip_partials || d->n_nodes < 2 || d->n_states < 1 || d->n_patterns < 1 ||
        d->n_rates < 1)
        return fail(MCP_ERR_INVALID, "plan: bad sizes or null pointers");
    const int N = d->n_nodes, S = d->n_states, C = d->n_patterns;
    std::vector<char> has_child(N, 0);
    for (int v = 0; v < N; ++v) {
        const int p = d->parent[v];
        if (p < 0 || p > N) return fail(MCP_ERR_INVALID, "parent num out of range");
        if (p > 0) has_child[p - 1] = 1;
    }
    std::vector<int32_t> leaves;
    for (int v = 0; v < N; ++v)
        if (!has_child[v]) leaves.push_back(v + 1);
    std::vector<double> tips((size_t)S * C * leaves.size());
    for (size_t j = 0; j < leaves.size(); ++j)
        std::memcpy(tips.data() + (size_t)S * C * j,
                    d->tip_partials + (size_t)S * C * (leaves[j] - 1), sizeof(double) * S * C);
    mcp_dataset_desc ds;
    ds.n_states = S;
    ds.n_patterns = C;
    ds.n_leaves = (int32_t)leaves.size();
    ds.leaf_nums = leaves.data();
    ds.tip_partials = tips.data();
    ds.pattern_weights = d->pattern_weights;
    ds.flags = d->flags;
    ds.tip_codes = nullptr;
    ds.n_codes = 0;
    ds.code_table = nullptr;
    mcp_item_desc it;
    it.dataset = 0;
    it.n_nodes = N;
    it.n_rates = d->n_rates;
    it.parent = d->parent;
    mcp_batch_desc bd;
    bd.n_datasets = 1;
    bd.datasets = &ds;
    bd.n_items = 1;
    bd.items = &it;
    bd.device = d->device;
    bd.waves_per_block = 0;
    return engine_create(out, &bd);
}

void mcp_plan_destroy(mcp_plan* p) { delete p; }

int mcp_set_model(mcp_plan* p, const double* U, const double* D, const double* Uinv, double mu,
                  const double* rates, const double* pi) {
    return engine_set_model(p, 0, U, D, Uinv, mu, rates, pi);
}

int mcp_set_topology(mcp_plan* p, const int32_t* parent) { return engine_set_topology(p, 0, parent); }

int mcp_logl(mcp_plan* p, const double* blv, double* ll) {
    if (p && p->items.size() != 1) return fail(MCP_ERR_INVALID, "handle is a batch");
    return engine_eval(p, blv, ll, nullptr, false);
}

int mcp_logl_grad(mcp_plan* p, const double* blv, double* ll, double* grad) {
    if (p && p->items.size() != 1) return fail(MCP_ERR_INVALID, "handle is a batch");
    if (!grad) return fail(MCP_ERR_INVALID, "null gradient");
    return engine_eval(p, blv, ll, grad, true);
}

int mcp_batch_create(mcp_batch** out, const mcp_batch_desc* d) { return engine_create(out, d); }

void mcp_batch_destroy(mcp_batch* b) { delete b; }

int mcp_batch_set_model(mcp_batch* b, int32_t item, const double* U, const double* D,
                        const double* Uinv, double mu, const double* rates, const double* pi) {
    return engine_set_model(b, item, U, D, Uinv, mu, rates, pi);
}

int mcp_batch_set_topology(mcp_batch* b, int32_t item, const int32_t* parent) {
    return engine_set_topology(b, item, parent);
}

int mcp_batch_logl(mcp_batch* b, const double* blv, double* ll) {
    return engine_eval(b, blv, ll, nullptr, false);
}

int mcp_batch_logl_grad(mcp_batch* b, const double* blv, double* ll, double* grad) {
    if (!grad) return fail(MCP_ERR_INVALID, "null gradient");
    return engine_eval(b, blv, ll, grad, true);
}

int mcp_hmc_set_prior(mcp_batch* b, int32_t item, int32_t kind, double scale) {
    if (!b) return fail(MCP_ERR_INVALID, "null handle");
    if (item < 0 || item >= (int32_t)b->items.size()) return fail(MCP_ERR_INVALID, "item out of range");
    if (kind != 0 && kind != 1) return fail(MCP_ERR_INVALID, "unknown prior");
    if (kind == 1 && !(scale > 0.0)) return fail(MCP_ERR_INVALID, "prior scale must be positive");
    DeviceGuard guard(b->device);
    HIP_TRY(hipStreamSynchronize(b->stream));
    int st = hmc_prepare(b);
    if (st) return st;
    b->hmc_items[item].prior = kind;
    b->hmc_items[item].prior_scale = kind == 1 ? scale : 1.0;
    HIP_TRY(hipMemcpy(b->dhmc_items.as<HmcItem>() + item, &b->hmc_items[item], sizeof(HmcItem),
                      hipMemcpyHostToDevice));
    return MCP_OK;
}

int mcp_logl_grad_by_rate(mcp_plan* p, const double* blv, double* ll, double* grad_vr) {
    return engine_eval_by_rate(p, blv, ll, grad_vr);
}

int mcp_logl_grad_model(mcp_plan* p, const double* blv, double* ll, double* grad, double* dlogl_dP,
                        double* dlogl_dpi) {
    return engine_eval_model(p, blv, ll, grad, dlogl_dP, dlogl_dpi);
}

int mcp_batch_logl_grad_model(mcp_batch* b, const double* blv_concat, double* ll,
                              double* grad_concat, double* dlogl_dP_concat,
                              double* dlogl_dpi_concat) {
    return engine_eval_model(b, blv_concat, ll, grad_concat, dlogl_dP_concat, dlogl_dpi_concat);
}

int mcp_batch_logl_grad_by_rate(mcp_batch* b, const double* blv_concat, double* ll,
                                double* grad_vr_concat) {
    return engine_eval_by_rate(b, blv_concat, ll, grad_vr_concat);
}

int mcp_hmc_set_prior_params(mcp_batch* b, int32_t item, int32_t kind, const double* params,
                             int32_t nparams) {
    if (!b) return fail(MCP_ERR_INVALID, "null handle");
    if (item < 0 || item >= (int32_t)b->items.size()) return fail(MCP_ERR_INVALID, "item out of range");
    const int need = kind == 0 ? 0 : kind == 1 ? 1 : kind == 2 ? 4 : -1;
    if (need < 0) return fail(MCP_ERR_INVALID, "unknown prior");
    if (nparams != need || (need > 0 && !params))
        return fail(MCP_ERR_INVALID, "wrong number of prior parameters");
    for (int k = 0; k < need; ++k)
        if (!(params[k] > 0.0) || !std::isfinite(params[k]))
            return fail(MCP_ERR_INVALID, "prior parameters must be positive and finite");
    DeviceGuard guard(b->device);
    HIP_TRY(hipStreamSynchronize(b->stream));
    int st = hmc_prepare(b);
    if (st) return st;
    HmcItem& h = b->hmc_items[item];
    h.prior = kind;
    h.prior_scale = kind == 1 ? params[0] : 1.0;
    for (int k = 0; k < 4; ++k) h.p[k] = (kind == 2) ? params[k] : 0.0;
    HIP_TRY(hipMemcpy(b->dhmc_items.as<HmcItem>() + item, &h, sizeof(HmcItem),
                      hipMemcpyHostToDevice));
    return MCP_OK;
}

int mcp_hmc_leapfrog(mcp_batch* b, const double* eps_delta, double* x, double* r, double* g,
                     double* logpost, double* ll) {
    return engine_leapfrog(b, eps_delta, x, r, g, logpost, ll);
}

int mcp_use_graph(mcp_batch* b, int32_t on) {
    if (!b) return fail(MCP_ERR_INVALID, "null handle");
    DeviceGuard guard(b->device);
    if (b->stream) HIP_TRY(hipStreamSynchronize(b->stream));
    b->use_graph = on != 0;
    if (!b->use_graph) b->drop_graphs();
    return MCP_OK;
}

int mcp_set_schedule(mcp_batch* b, int32_t mode) {
    if (!b) return fail(MCP_ERR_INVALID, "null handle");
    if (mode != MCP_SCHEDULE_FUSED && mode != MCP_SCHEDULE_LEVELS && mode != MCP_SCHEDULE_HYBRID)
        return fail(MCP_ERR_INVALID, "unknown schedule");
    DeviceGuard guard(b->device);
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->schedule == mode) return MCP_OK;
    b->schedule = mode;
    return rebuild_schedule(b);
}

int mcp_set_hybrid_width(mcp_batch* b, int32_t width) {
    if (!b) return fail(MCP_ERR_INVALID, "null handle");
    if (width < 1) return fail(MCP_ERR_INVALID, "hybrid width must be >= 1");
    DeviceGuard guard(b->device);
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->hybrid_width == width) return MCP_OK;
    b->hybrid_width = width;
    return b->schedule == MCP_SCHEDULE_HYBRID ? rebuild_schedule(b) : MCP_OK;
}

int mcp_set_fused_waves(mcp_batch* b, int32_t waves) {
    if (!b) return fail(MCP_ERR_INVALID, "null handle");
    if (waves < 1 || waves > 8) return fail(MCP_ERR_INVALID, "fused waves must be 1..8");
    DeviceGuard guard(b->device);
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->waves = waves;
    b->drop_graphs();  // captured fused launches hold the old block size
    // the fused schedule groups its steps by the block's wavefront count
    if (b->schedule == MCP_SCHEDULE_FUSED) return rebuild_schedule(b);
    return MCP_OK;
}

int mcp_set_tiles(mcp_batch* b, int32_t S, int32_t cpt) {
    if (!b) return fail(MCP_ERR_INVALID, "null handle");
    if (!supported_states(S)) return fail(MCP_ERR_UNSUPPORTED, "no kernels for this state count");
    if (cpt < 0) cpt = default_cpt(S);
    if (!cpt_available(S, cpt))
        return fail(MCP_ERR_UNSUPPORTED, "no kernels for this (state count, columns per thread)");
    DeviceGuard guard(b->device);
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->cpt(S) == cpt) return MCP_OK;
    b->cpt_sel[S] = cpt;
    return relayout_tiles(b);
}

int mcp_get_tiles(mcp_batch* b, int32_t S, int32_t* cpt) {
    if (!b || !cpt) return fail(MCP_ERR_INVALID, "null argument");
    if (!supported_states(S)) return fail(MCP_ERR_UNSUPPORTED, "no kernels for this state count");
    *cpt = b->cpt(S);
    return MCP_OK;
}

int mcp_autotune(mcp_batch* b, int32_t reps, double* report, int32_t cap, int32_t* n_report) {
    if (!b) return fail(MCP_ERR_INVALID, "null handle");
    if (reps < 1 || reps > 1000) return fail(MCP_ERR_INVALID, "reps must be in 1..1000");
    for (auto& it : b->items)
        if (!it.model_set) return fail(MCP_ERR_STATE, "substitution model not set");
    DeviceGuard guard(b->device);
    HIP_TRY(hipStreamSynchronize(b->stream));
    std::vector<double> blv((size_t)b->total_blv, 0.1);
    if (b->evaluated) std::memcpy(blv.data(), b->hblv.p, blv.size() * sizeof(double));
    std::vector<double> ll(b->items.size()), grad((size_t)b->total_blv);
    std::vector<int> svals;
    for (auto& it : b->items)
        if (std::find(svals.begin(), svals.end(), it.S) == svals.end()) svals.push_back(it.S);
    std::sort(svals.begin(), svals.end());
    const bool graph = b->use_graph, timing = b->timing;
    b->use_graph = false;
    b->timing = false;
    int nrep = 0, st = MCP_OK;
    for (int S : svals) {
        std::vector<int> cands;
        for (int c = 0; c <= 8; ++c)
            if (cpt_available(S, c)) cands.push_back(c);
        int best = b->cpt(S);
        double best_ms = 1e300;
        for (int c : cands) {
            b->cpt_sel[S] = c;
            if ((st = relayout_tiles(b))) break;
            // one untimed evaluation, then the median of reps timed ones
            if ((st = engine_eval(b, blv.data(), ll.data(), grad.data(), true))) break;
            std::vector<double> t(reps);
            for (int k = 0; k < reps && !st; ++k) {
                const auto t0 = std::chrono::steady_clock::now();
                st = engine_eval(b, blv.data(), ll.data(), grad.data(), true);
                t[k] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                           .count();
            }
            if (st) break;
            std::sort(t.begin(), t.end());
            const double ms = t[reps / 2];
            if (report && nrep < cap) {
                report[3 * nrep] = S;
                report[3 * nrep + 1] = c;
                report[3 * nrep + 2] = ms;
            }
            ++nrep;
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
        if (st) break;
        b->cpt_sel[S] = best;
    }
    const int st2 = relayout_tiles(b);
    b->use_graph = graph;
    b->timing = timing;
    if (n_report) *n_report = nrep;
    return st ? st : st2;
}

int mcp_autotune_schedule(mcp_batch* b, int32_t reps, double* report, int32_t cap,
                          int32_t* n_report) {
    if (!b) return fail(MCP_ERR_INVALID, "null handle");
    if (reps < 1 || reps > 1000) return fail(MCP_ERR_INVALID, "reps must be in 1..1000");
    for (auto& it : b->items)
        if (!it.model_set) return fail(MCP_ERR_STATE, "substitution model not set");
    DeviceGuard guard(b->device);
    HIP_TRY(hipStreamSynchronize(b->stream));
    std::vector<double> blv((size_t)b->total_blv, 0.1);
    if (b->evaluated) std::memcpy(blv.data(), b->hblv.p, blv.size() * sizeof(double));
    std::vector<double> ll(b->items.size()), grad((size_t)b->total_blv);
    const bool graph = b->use_graph, timing = b->timing;
    b->use_graph = false;
    b->timing = false;
    // candidates: (schedule, hybrid width)
    const int cands[][2] = {{MCP_SCHEDULE_LEVELS, 0}, {MCP_SCHEDULE_HYBRID, 1},
                            {MCP_SCHEDULE_HYBRID, 2}, {MCP_SCHEDULE_HYBRID, 4},
                            {MCP_SCHEDULE_HYBRID, 8}, {MCP_SCHEDULE_HYBRID, 16},
                            {MCP_SCHEDULE_FUSED, 0}};
    const int keep_w = b->hybrid_width;
    int best_mode = b->schedule, best_w = keep_w, nrep = 0, st = MCP_OK;
    double best_ms = 1e300;
    for (const auto& c : cands) {
        b->schedule = c[0];
        b->hybrid_width = c[0] == MCP_SCHEDULE_HYBRID ? c[1] : keep_w;
        if ((st = rebuild_schedule(b))) break;
        if ((st = engine_eval(b, blv.data(), ll.data(), grad.data(), true))) break;
        std::vector<double> t(reps);
        for (int k = 0; k < reps && !st; ++k) {
            const auto t0 = std::chrono::steady_clock::now();
            st = engine_eval(b, blv.data(), ll.data(), grad.data(), true);
            t[k] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                       .count();
        }
        if (st) break;
        std::sort(t.begin(), t.end());
        const double ms = t[reps / 2];
        if (report && nrep < cap) {
            report[3 * nrep] = c[0];
            report[3 * nrep + 1] = c[1];
            report[3 * nrep + 2] = ms;
        }
        ++nrep;
        if (ms < best_ms) {
            best_ms = ms;
            best_mode = c[0];
            best_w = b->hybrid_width;
        }
    }
    b->schedule = best_mode;
    b->hybrid_width = best_w;
    const int st2 = rebuild_schedule(b);
    b->use_graph = graph;
    b->timing = timing;
    if (n_report) *n_report = nrep;
    return st ? st : st2;
}

int mcp_get_schedule(mcp_batch* b, int32_t* mode, int32_t* hybrid_width) {
    if (!b || !mode) return fail(MCP_ERR_INVALID, "null argument");
    *mode = b->schedule;
    if (hybrid_width) *hybrid_width = b->hybrid_width;
    return MCP_OK;
}

int mcp_set_timing(mcp_batch* b, int32_t on) {
    if (!b) return fail(MCP_ERR_INVALID, "null handle");
    DeviceGuard guard(b->device);
    if (on && !b->ev_begin) {
        HIP_TRY(hipEventCreate(&b->ev_begin));
        HIP_TRY(hipEventCreate(&b->ev_end));
    }
    b->timing = on != 0;
    return MCP_OK;
}

int mcp_get_timing(mcp_batch* b, double* ms5) {
    if (!b || !ms5) return fail(MCP_ERR_INVALID, "null argument");
    for (int k = 0; k < 5; ++k) ms5[k] = b->last_ms[k];
    return MCP_OK;
}

int mcp_get_launch_times(mcp_batch* b, double* ms, int32_t* phase, int32_t cap) {
    if (!b) return -fail(MCP_ERR_INVALID, "null handle");
    const int n = std::min<int>(cap, (int)b->kms.size());
    for (int k = 0; k < n; ++k) {
        if (ms) ms[k] = b->kms[k];
        if (phase) phase[k] = b->kphase[k];
    }
    return (int)b->kms.size();
}

int mcp_get_launch_bytes(mcp_batch* b, double* bytes, int32_t cap) {
    if (!b) return -fail(MCP_ERR_INVALID, "null handle");
    const int n = std::min<int>(cap, (int)b->kms.size());
    for (int k = 0; k < n && bytes; ++k) bytes[k] = k < (int)b->kbytes.size() ? b->kbytes[k] : 0.0;
    return (int)b->kms.size();
}

int mcp_get_bytes(mcp_batch* b, double* bytes2) {
    if (!b || !bytes2) return fail(MCP_ERR_INVALID, "null argument");
    bytes2[0] = b->bytes_fwd;
    bytes2[1] = b->bytes_adj;
    return MCP_OK;
}

int mcp_get_sizes(mcp_batch* b, int64_t* sizes2) {
    if (!b || !sizes2) return fail(MCP_ERR_INVALID, "null argument");
    sizes2[0] = (int64_t)b->items.size();
    sizes2[1] = (int64_t)b->total_patterns;
    return MCP_OK;
}

}  // extern "C"
