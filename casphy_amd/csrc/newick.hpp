// newick.hpp -- Newick trees numbered as ape's read.tree numbers them (host side).
#pragma once

#include <string>
#include <vector>

#include "host.hpp"

namespace casphy {

struct NewickTree {
    u32 ntip = 0, nnode = 0;
    std::vector<int32_t> anc, des;    // edge matrix columns, 1-based, cladewise (preorder) order
    std::vector<std::string> labels;  // tip labels by tip number (tip 1 first)
};

// Reads the tree that starts at text[pos] (leading blanks and [comments]
// skipped) up to its ';', and moves pos past it; ntip = 0 when only blanks
// remain.  With `order` (tip labels), tip i is numbered by the position of
// its label there (every label once) instead of by order of appearance.
NewickTree read_newick(const char *text, u64 len, u64 &pos, const std::vector<std::string> *order = nullptr);

// Transfer bootstrap expectation of the reference tree's clades from the
// transfer minima of casphy_transfer_trees.
void tbe_support(const u32 *ymin, u32 nrep, u32 ny, const u32 *ysize, u32 ntip, double *support);

}  // namespace casphy
