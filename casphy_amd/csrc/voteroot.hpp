// voteroot.hpp -- root votes of outgroup samples over a tree's edge states.
#pragma once

#include <vector>

#include "common.hpp"

namespace casphy {

// Candidate order of the vote: candidates[r] is the node whose count competes
// with rank r (ties go to the smallest rank); parent[v] is v's parent edge's
// upper node (~0u: v is no edge's child, its edge state is zero).
struct RootVote {
    std::vector<u32> candidates;
    std::vector<u32> parent;
};

// voteRoot2's order (every node; the reference's block-of-8 and 32-lane
// reductions) or findRoot's (edge children in edge order).  edge: nedge x 2
// 1-based (parent, child) ids in 1..nnode.
RootVote plan_vote_voteroot2(const u32 *edge, u32 nedge, u32 nnode);
RootVote plan_vote_findroot(const u32 *edge, u32 nedge, u32 nnode);

// Per sample row: (count << 32) | (0xFFFFFFFF - rank) of the winning candidate
// (0 when there is none).  state: host nnode x nbin words of byte codes,
// samples: host nsample x nbin.
void vote_root_keys(const RootVote &plan, const u32 *state, u32 nnode, u32 nbin, const u32 *samples, u32 nsample,
                    u64 *keys, hipStream_t s);

}  // namespace casphy
