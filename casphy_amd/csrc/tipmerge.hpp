// tipmerge.hpp -- tipMerge over device-resident neighbour lists (declarations).
#pragma once

#include "common.hpp"

namespace casphy {

// Cluster labels of cu/tipMerge.cu for n tips whose ascending neighbour lists
// (ids < n, no repeats) sit on the device as off[n + 1] (u64) / ids (u32).
// bits (may be null): the same lists as bit rows, bit t of row j set iff t is
// in the list of j, rows of bstride words (a radius mask); the dense form
// builds them when null.  member: n host int32 labels.  Synchronises s.
void tip_merge_device(const u64 *off, const u32 *ids, u32 n, const u32 *bits, u32 bstride, double p,
                      int32_t *member, hipStream_t s);

// The reference's limit on one list (MAXNEIGHBOR, cu/tipMerge.cu:7,183-186).
constexpr u32 TIP_MERGE_MAX_LIST = 1024u * 1024u;

}  // namespace casphy
