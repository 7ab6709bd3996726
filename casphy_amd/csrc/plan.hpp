// plan.hpp -- host-side planning of the tree kernels (no HIP: these sources
// also build with a plain C++ compiler, so the CPU tests run them under the
// address and undefined-behaviour sanitizers).
#pragma once

#include <vector>

#include "host.hpp"

namespace casphy {

// ------------------------------------------------------------ clade rows
// What the rows of nodeDesBig are made of: the tips in depth-first order, and
// per written row (in nodeDesBig's order) the run of that order below its node.
struct CladePlan {
    u32 nbin = 0;                  // words per row, ceil(ntip / 32)
    std::vector<u32> order;        // tip ids (0-based) in walk order
    std::vector<u32> lo, cnt;      // per row: run start and length
    std::vector<int32_t> size;     // nodeDesBig's return vector (nnode - 1)
};

// anc/des: R's 1-based edge matrix columns (tips 1..ntip, internal nodes
// ntip+1..ntip+nnode, one root, every node's edge before its children's).
CladePlan plan_clades(const int32_t *anc, const int32_t *des, u32 nedge, u32 nnode, u32 ntip);

// ------------------------------------------------------------ tree walks
// Op words of a transfer-distance walk (see tbe.hip): flag bits over a
// 29-bit argument.
//   TIP p  = POS | STACK | p   push the running value, start at [p in run]
//   ACC p  = POS | p           add [p in run]
//   ADD    = STACK [| ROW | s] add the value under the top, pop [, row of size s]
//   ROW s  = ROW | s           the running value is the count of a clade row of s tips
constexpr u32 OP_POS = 1u << 31, OP_STACK = 1u << 30, OP_ROW = 1u << 29;
constexpr u32 ARG_MASK = OP_ROW - 1;   // tip positions and clade sizes below 2^29
constexpr u32 NOT_IN_TREE = ARG_MASK;  // position of a tip the lane tree lacks
constexpr u32 WALK_STACK_MAX = 32;     // values per lane: trees up to 2^31 tips
constexpr u32 OP_BLOCK = 16;           // programs are padded to a multiple of this

inline u32 op_tip(u32 p) { return OP_POS | OP_STACK | p; }
inline u32 op_acc(u32 p) { return OP_POS | p; }
inline u32 op_add() { return OP_STACK; }
inline u32 op_row(u32 size) { return OP_ROW | size; }

// Children lists (0-based nodes: tips 0..ntip-1, internal ntip..), tips below
// every node and the clade-row nodes (internal, not the root) of a cladewise
// edge matrix that plan_clades has accepted.
struct TreeShape {
    u32 ntip = 0, nnode = 0, root = 0;
    std::vector<u32> kid_off, kid;   // children of node v: kid[kid_off[v] .. kid_off[v+1])
    std::vector<u32> tips;           // tips below each node
    std::vector<unsigned char> is_row;
};
TreeShape tree_shape(const int32_t *anc, const int32_t *des, u32 nedge, u32 nnode, u32 ntip);

// Walk of a tree, heaviest child first: op words with tip ids as arguments
// (mapped to positions of the lane tree per pairing) and the deepest stack
// the walk needs.
struct WalkSkeleton {
    std::vector<u32> ops;  // TIP / ACC arg = tip id
    u32 depth = 0;
};
WalkSkeleton walk_skeleton(const TreeShape &t);

// Appends a walk with tip ids mapped to positions of the lane tree, padded to
// a multiple of OP_BLOCK words with ACC ops of a position no run contains.
void emit_walk(const WalkSkeleton &w, const std::vector<u32> &pos, std::vector<u32> &ops);
inline u32 padded_ops(const WalkSkeleton &w) { return (u32)((w.ops.size() + OP_BLOCK - 1) / OP_BLOCK * OP_BLOCK); }

// Per-tree data of the transfer walks: the clade plan (rows in nodeDesBig's
// order, the run of each row in the depth-first tip order), the position of
// every tip in that order, and the tree's walk.
struct TreeData {
    CladePlan plan;
    std::vector<u32> pos;
    WalkSkeleton walk;
};
TreeData tree_data(const int32_t *anc, const int32_t *des, u32 nedge, u32 nnode, u32 ntip);

// ------------------------------------------------------------ nodeKeep
// cu/nodeKeepBit.cu:186-211: edges (0-based ids, tips 0..ntip-1) walked in
// order; an edge to internal node d is skipped while the last evaluated node
// was kept and the edge leaves a node whose id is not below it, otherwise d is
// evaluated (retain[d - ntip]).  Writes the kept ids, returns their count.
u32 node_keep_walk(const u32 *edge, u32 nedge, const u32 *retain, u32 ntip, u32 *kept);

}  // namespace casphy
