// clades.hpp -- clade bitsets of a rooted tree (declarations).
#pragma once

#include <vector>

#include "common.hpp"

namespace casphy {

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
// rows: device buffer of plan.lo.size() rows of `stride` >= plan.nbin words
// (words past nbin are zeroed).
void build_clade_rows(const CladePlan &p, u32 *rows, u32 stride, hipStream_t s);

}  // namespace casphy
