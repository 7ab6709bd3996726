// tiles.hpp -- exact integer pair-scoring tiles for gfx950 (declarations).
//
// One kernel family scores every (A row, B row) pair of a tile grid with an
// exact integer metric and hands the 8x8 per-thread block of scores to a fused
// epilogue.  Metrics:
//   AND   popcount(a & b)                       binary / one-hot indel score
//   XOR   popcount(a ^ b)                       transfer-distance Hamming
//   DNA2  popcount(ma & mb & ~((ah^bh)|(al^bl))) 2-bit ACGT + non-gap mask,
//         three words (hi, lo, mask) per 32 positions
//   NIB   count of non-zero nibbles of (a & b)  IUPAC nibble codes, 8 per word
// Weighted scores (non-negative integer weights) run as bit planes: the A row
// is masked with the sites whose weight has bit p set and the accumulators are
// shifted between planes (Horner), so the sum stays exact in int32.
#pragma once

#include <string>

#include "common.hpp"

namespace casphy {

enum Metric : int { M_AND = 0, M_XOR = 1, M_DNA2 = 2, M_NIB = 3 };
enum Epi : int { E_MASK = 0, E_SCORES = 1, E_HIST = 2, E_LOCALMAX = 3, E_TBE = 4, E_PAIRS = 5, E_ARGMAX = 6, E_COUNTS = 7 };

constexpr int TILE = 128;      // rows per tile side
constexpr int KC_BIT = 32;     // words per LDS stage (AND / XOR / NIB)
constexpr int KC_DNA2 = 24;    // 8 groups of (hi, lo, mask)
constexpr int MAX_PLANES = 24;
constexpr int MAX_RADII = 16;

struct TileArgs {
    const u32 *A = nullptr;      // A rows, ldw words each
    const u32 *Aidx = nullptr;   // optional gather of A rows
    u32 nA = 0;
    const u32 *B = nullptr;
    u32 nB = 0;
    u32 ldw = 0;                 // row stride in words (multiple of the stage width)
    u32 kw = 0;                  // words per plane actually scored (<= ldw, multiple of stage)
    u32 ntA = 0, ntB = 0;        // tiles along A and B
    u32 ntiles = 0;              // tiles in the launch
    int triangle = 0;            // 1: A == B, only tiles I <= J
    // weight planes (AND metric only); nplanes == 0 means unit weights
    u32 nplanes = 0;
    const u32 *pmask = nullptr;  // nplanes x kw site masks
    u32 pshift[MAX_PLANES] = {};  // shift applied before plane p (p >= 1)
    u32 ptail = 0;                // shift applied after the last plane
    // epilogue parameters
    int32_t thr = 0;             // E_MASK
    const int32_t *thr_dev = nullptr;  // E_MASK: threshold read from device memory when set
    u32 *mask = nullptr;
    u32 mstride = 0;
    int32_t *S = nullptr;        // E_SCORES
    u64 ldS = 0;
    u32 nradii = 0;              // E_HIST
    int32_t radii[MAX_RADII] = {};
    u32 *counts = nullptr;       // nradii x nA
    u64 *keys = nullptr;         // E_LOCALMAX
    u32 ntip = 0;                // E_TBE
    u32 *amin = nullptr, *bmin = nullptr;
};

// Launches the tile kernel for (metric, epilogue) on `stream`; the tile shape
// and stage width come from the autotuner (or CASPHY_VALU_VARIANT).  The
// caller's tile grid (ntA, ntB, ntiles) is in 128 x 128 tiles.
void launch_tiles(int metric, int epi, const TileArgs &a, hipStream_t stream);

// JSON list of the VALU tile autotuner's choices so far (one entry per shape
// class, with the time of every candidate).
std::string valu_tuning_json();

// Stage width (words) of a metric: rows must be padded to a multiple of it.
inline u32 stage_words(int metric) { return metric == M_DNA2 ? KC_DNA2 : KC_BIT; }

}  // namespace casphy
