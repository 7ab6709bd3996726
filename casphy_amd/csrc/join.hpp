// join.hpp -- greedy joins whose input rows are already on the device (the
// request loop's resident matrices), unit weights.
#pragma once

#include <vector>

#include "common.hpp"

namespace casphy {

// raw: n x W cell-major bit rows; edge_host: (2n-2) x 2 raw edges.
void join_bit_dev(const u32 *raw, u32 n, u32 W, int ancestral_type, u32 *edge_host, hipStream_t s);
// raw: n x ncol allele codes (-1 missing, 0 unmutated).
void join_indel_dev(const int32_t *raw, u32 n, u32 ncol, u32 *edge_host, hipStream_t s);
// Upper bound on the workgroups of a one-launch join step for joins started
// from the calling host thread (0: the single-join rule); the request loop's
// workers lower it so that concurrent joins share the device.
void set_join_workgroup_cap(u32 cap);
// Many joins in one persistent launch: job k is rows offs[k] .. offs[k+1] of raw.
void join_bit_batch_dev(const u32 *raw, const std::vector<u32> &offs, u32 W, int ancestral_type, u32 *edge_host,
                        hipStream_t s);
void join_indel_batch_dev(const int32_t *raw, const std::vector<u32> &offs, u32 ncol, u32 *edge_host,
                          hipStream_t s);

}  // namespace casphy
