// mfma.hpp -- exact AND-count tiles on the int8 matrix cores (declarations).
#pragma once

#include "tiles.hpp"

namespace casphy {

struct MfmaArgs {
    const u32 *P = nullptr;  // stage-major packed bits: [stages][nrows_pad][4]
    u32 nrows_pad = 0;
    u32 stages = 0;          // 128 sites each
    u32 nA = 0, nB = 0;      // valid rows (A == B: self scoring, upper-triangle tiles)
    u32 ntiles = 0;
    int32_t thr = 0;         // E_MASK
    u32 *mask = nullptr;
    u32 mstride = 0;
    int32_t *S = nullptr;    // E_SCORES
    u64 ldS = 0;
};

// cell-major [n][ldw] -> stage-major P (rows padded to a multiple of 128).
void launch_to_stage_major(const u32 *mat, u32 n, u32 ldw, u32 *P, hipStream_t s);
void launch_mfma_tiles(int epi, const MfmaArgs &a, hipStream_t s);

}  // namespace casphy
