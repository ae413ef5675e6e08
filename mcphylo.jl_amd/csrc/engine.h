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
//   with the pattern axis padded to Cp, a multiple of the task width, so
//   every task's columns are one contiguous span.
#pragma once
#include <stdint.h>

namespace mcp {

// A child of a level job's node, with every address resolved on the host so
// a task reaches its operands after one round of scalar loads.
struct ChildRef {
    const double* F;       // internal child: partial slab at r = 0 ([R][Cp][S]); leaf: nullptr
    const uint8_t* codes;  // leaf child: its tip-code row ([Cp]); internal: nullptr
    const double* P;       // internal: P of the child's branch at r = 0 ([R][S][S], P[s][s1]);
                           // leaf: P . tipvec rows at r = 0 ([R][K][S])
    const double* dP;      // same layouts for dP/dt (pre-order pass)
    double* G;             // internal child: its pre-order slab at r = 0; leaf: nullptr
    int32_t gpart;         // offset of the branch's gradient partial sums
    int32_t mpart;         // offset of the branch's d logL / dP partial sums (model-gradient
                           // buffer, layout [r][S*S][tile]), -1 when not laid out
    int32_t leaf;          // leaf child: its row of the dataset's code matrix; internal: -1
    int16_t gslot;         // internal child under a fused LDS slab stage: slot of its
                           // pre-order slab, else -1
    int16_t fslot;         // the same: slot of its partial's post-order copy, else -1
};

constexpr int kInlineChildren = 3;

// One internal node of one item.  Post-order launches read F, lpart,
// is_root; pre-order launches read G (the node's own pre-order slab) and the
// children's G/gpart.  A job expands into R * ntiles tasks.
struct LevelJob {
    int32_t S, C, Cp, R, K, ntiles;
    int32_t nchild;
    int32_t lpart;          // offset of this node's log-scaler partial sums
    int32_t is_root;
    int32_t rpart;          // root: offset of the d logL / d pi partial sums (model-gradient
                            // buffer, layout [S][r][tile]), else -1
    int32_t gslot;          // fused LDS slab stage: slot of the own pre-order slab (non-root),
                            // else -1
    int32_t fslot;          // the same: slot of the own partial's post-order copy, else -1
    const double* pi;       // [S]
    double* F;              // output partial slab at r = 0 (non-root), else nullptr
    const double* G;        // own pre-order slab at r = 0 (non-root), else nullptr (root: pi)
    const double* weights;  // [Cp], zero in the padding
    ChildRef kid[kInlineChildren];
    const double* tiptab;   // [K][S] tip vectors of the state codes
    const ChildRef* more;   // children beyond kInlineChildren
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
    double* P;                 // [N-1][R][S][S]
    double* dP;
    double* PT;                // [n_leaves][R][K][S]
    double* dPT;
};

// Per-item record of the fused schedule (one launch per state count).
struct ItemSched {
    int32_t block_off;     // first block of this item in the launch
    int32_t R, ntiles;
    int32_t lf0, lf1;      // post-order levels: [lf0, lf1) into lvl_f (lvl_f[k] = first job)
    int32_t la0, la1;      // pre-order levels:  [la0, la1) into lvl_a
    int32_t lds_leaves;    // code rows each block stages in LDS (0: leaves read global codes)
    const uint8_t* codes;  // the item's dataset code matrix [leaf][Cp]
    int32_t Cp;
    int32_t lds_g;         // byte offset of the block's LDS slab stage (slot q: [S][64 * CPT]
                           // doubles at lds_g + q * 512 * S * CPT), holding the post-order
                           // frontier's partials, then the pre-order slabs; -1: none
};

// Reduction descriptor: out[dst] = sum(parts[start .. start+len)).
// One fixed-order sum: out[dst] (or tmp[dst] when to_tmp) = sum of
// src[start .. start+len).  Long sums are split into chunks summed into tmp
// by a first launch and combined by a second, in a fixed order.
struct RedOp {
    int32_t dst;
    int32_t start;
    int32_t len;
    int32_t to_tmp;
};

}  // namespace mcp

namespace mcp {

// Per-item record of the device leapfrog (reference: refraction!,
// src/samplers/tree_hamiltonian/refraction.jl:2-31, without NNI moves).
// State buffers are packed per item at offset `off` (n = N-1 branches).
// Branch-length prior: 0 none; 1 exponentialBL(prior_scale)
// (src/Likelihood/Prior.jl:50-56,80-83); 2 CompoundDirichlet with
// p = (alpha, a, beta, c) (Prior.jl:1-37).
struct HmcItem {
    int32_t n;             // branches
    int32_t off;           // offset of this item's branch vectors
    int32_t out;           // offset of (ll, grad) in the engine's output buffer
    int32_t prior;         // prior kind (above)
    double prior_scale;
    double p[4];
};

}  // namespace mcp
