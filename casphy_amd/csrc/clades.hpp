// clades.hpp -- clade bitsets of a rooted tree (declarations).
#pragma once

#include <vector>

#include "common.hpp"
#include "plan.hpp"

namespace casphy {

// rows: device buffer of plan.lo.size() rows of `stride` >= plan.nbin words
// (words past nbin are zeroed).
void build_clade_rows(const CladePlan &p, u32 *rows, u32 stride, hipStream_t s);

}  // namespace casphy
