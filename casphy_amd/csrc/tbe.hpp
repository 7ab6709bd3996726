// tbe.hpp -- transfer distances between the clades of whole trees (declarations).
#pragma once

#include "common.hpp"

namespace casphy {

// Transfer minima of the reference tree's clade rows against nrep replicate
// trees (R edge-matrix columns, cladewise; replicate b's edges at
// x_edge_offsets[b] .. x_edge_offsets[b+1]), rows in nodeDesBig's order:
// ymin[b * ny + y], ysize[ny], xmin / xsize per replicate row in replicate
// order.  The same values as the bitset tiles, by a walk of one tree per
// lane group; replicates stream through the device in chunks.
void transfer_trees_dp(const int32_t *y_anc, const int32_t *y_des, u32 y_nedge, u32 y_nnode, const int32_t *x_anc,
                       const int32_t *x_des, const u32 *x_edge_offsets, const u32 *x_nnode, u32 nrep, u32 ntip,
                       u32 *ymin, u32 *ysize, u32 *xmin, u32 *xsize, hipStream_t s);

}  // namespace casphy
