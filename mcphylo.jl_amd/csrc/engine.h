// Shared device/host structures of the pruning engine (gfx950).
//
// One engine evaluates a batch of "items".  An item is one (tree, chain):
// a rooted tree with its own branch lengths, substitution model and
// partial-likelihood slabs, bound to a "dataset" (tip states of one
// alignment, shared by every chain of the same tree).  A single PhyloDist
// evaluation is a batch of one item.
//
// HBM layout of a partial slab (one internal node of one item):
//   slab[r][c][s]  (state fastest, then pattern, then category) -- the
//   reference's own data[s, c, r, v] order (LikelihoodCalculator_Node.jl) --
//   with the pattern axis padded to Cp = roundup(C, 64) so every category
//   slab starts on a 64-pattern boundary and a wavefront's loads are one
//   contiguous span.
#pragma once
#include <stdint.h>

namespace mcp {

constexpr int kMaxStates = 64;
constexpr int kPadPatterns = 64;

// Per-node record of one item (index = num-1).
struct NodeInfo {
    int32_t first_child;   // offset into Item::children
    int32_t nchild;
    int32_t fslot;         // partial slab index (internal non-root), else -1
    int32_t leaf;          // row of the dataset's tip codes (leaves), else -1
    int32_t gslot;         // pre-order slab index in ring buffer gbuf (internal non-root), else -1
    int32_t gbuf;          // 0/1: which pre-order ring buffer (depth parity)
    int32_t lpart;         // offset of this node's log-scaler partial sums in parts[] (internal), else -1
    int32_t gpart;         // offset of this branch's gradient partial sums in parts[] (non-root), else -1
    int32_t is_root;
    int32_t pad_;
};

// Per-item device descriptor.
struct Item {
    int32_t S, C, Cp, R, K, N, ntiles, tile;   // tile == threads per task (patterns per task)
    const NodeInfo* nodes;     // [N]
    const int32_t* children;   // child indices (num-1)
    const uint8_t* codes;      // [n_leaves][Cp] tip codes
    const double* weights;     // [Cp] (zero in the padding)
    const double* pi;          // [S]
    const double* P;           // [N-1][R][S][S] row-major P[s][s1]
    const double* dP;          // [N-1][R][S][S]
    const double* PT;          // [n_leaves][R][K][S]  P . tipvec_k  (leaf branches)
    const double* dPT;         // [n_leaves][R][K][S]  dP . tipvec_k
    double* F;                 // [n_fslots][R][Cp][S]
    double* G0;                // pre-order ring buffer 0 [g0 slots][R][Cp][S]
    double* G1;                // pre-order ring buffer 1
};

// Work unit of one level launch: one node of one item; it expands to
// R * ntiles tasks (one per (category, pattern tile)).
struct Job {
    int32_t item;
    int32_t node;              // num-1
};

// Inputs of the transition-matrix builder, per item.
struct ModelDev {
    const double* U;           // [S][S] column-major (as given)
    const double* D;           // [S]
    const double* Uinv;        // [S][S] column-major
    const double* rates;       // [R]
    const double* blv;         // [N-1]
    const double* tiptab;      // [K][S]
    const int32_t* leaf_of;    // [N] leaf row or -1
    double mu;
    int32_t S, R, N, K;
    double* P;
    double* dP;
    double* PT;
    double* dPT;
};

// Reduction descriptor: out[dst] = sum(parts[start .. start+len)).
struct RedOp {
    int32_t dst;
    int32_t start;
    int32_t len;
    int32_t pad_;
};

}  // namespace mcp
