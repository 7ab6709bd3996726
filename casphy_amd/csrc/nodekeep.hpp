// nodekeep.hpp -- nodeKeep over a joined tree (declarations).
#pragma once

#include "common.hpp"

namespace casphy {

// Kept internal node ids, in edge order, of nodeKeepBit (indel = false: rows
// of `width` bit words) or nodeKeepIndel (rows of `width` u32 states).  edge:
// nedge x 2, 0-based, tips 0..nnode, each internal node at most once as a
// descendant; state: nnode x width; outmat: nout x width.  kept holds nnode.
u32 node_keep(bool indel, const u32 *edge, u32 nedge, const u32 *state, u32 nnode, u32 width, const u32 *outmat,
              u32 nout, u32 cutoff, u32 *kept, hipStream_t s);

}  // namespace casphy
