// mfma.hpp -- exact AND-count tiles on the int8 matrix cores (declarations).
#pragma once

#include <string>

#include "tiles.hpp"

namespace casphy {

// Rows of every stage-major copy are padded to a multiple of this.
constexpr u32 STAGE_PAD = 256;
inline u32 stage_rows(u32 n) { return ceil_div(n, STAGE_PAD) * STAGE_PAD; }

// One problem of a group-mode launch (many independent row sets in one
// launch): its A and B rows start at rows a0 / b0 (multiples of 128) of the
// gathered stage-major operands and number nA / nB; its output is the mask of
// nA rows of `ld` words at word offset `off` (E_MASK) or the nA x nB scores
// of row stride `ld` at offset `off` (E_SCORES).
struct GroupDesc {
    u32 a0 = 0, b0 = 0;
    u32 nA = 0, nB = 0;
    u64 off = 0;
    u32 ld = 0, pad_ = 0;
};

struct MfmaArgs {
    const u32 *PA = nullptr;  // stage-major packed bits of the A rows: [stages][nrows_padA][4]
    const u32 *PB = nullptr;  // and of the B rows (== PA for self scoring)
    u32 padA = 0, padB = 0;   // padded row counts (multiples of 128)
    u32 stages = 0;           // 128 sites each
    const u32 *EA = nullptr;  // optional fp4 expansions of PA / PB (launch_expand_fp4); the
    const u32 *EB = nullptr;  // fp4 variant expands what is not given
    u32 nA = 0, nB = 0;       // valid rows
    u32 ntA = 0, ntB = 0;     // set by the launcher (tile grid depends on the variant)
    u32 sup = 0;              // set by the launcher: super-tile edge in tiles (0: plain order)
    u32 ntiles = 0;           // set by the launcher
    int triangle = 1;         // A == B, tiles I <= J
    int32_t thr = 0;          // E_MASK
    const int32_t *thr_dev = nullptr;  // E_MASK: threshold read from device memory when set
    u32 *mask = nullptr;
    u32 mstride = 0;
    int32_t *S = nullptr;     // E_SCORES
    u64 ldS = 0;
    u64 *keys = nullptr;      // E_LOCALMAX (per A row), E_ARGMAX (per B row)
    u32 ntip = 0;             // E_TBE: H = popA + popB - 2 <a, b>
    const u32 *popA = nullptr, *popB = nullptr;
    u32 *amin = nullptr, *bmin = nullptr;
    const u32 *arow_map = nullptr;  // E_TBE batches: padded A row -> output row (~0u = padding)
    const u32 *tile_rep = nullptr;  // E_TBE batches: 128-row block of A -> replicate (bmin row block)
    u32 bstride = 0;                // E_TBE batches: bmin entries per replicate
    // E_PAIRS: for i < j with bit j set in row i of `mask` (stride mstride
    // words), pout[prow[i] + pq[i * mstride + j / 32] + (set bits of that word
    // below bit j)] = <a_i, b_j> (prow holds the row's base minus its bits <= i,
    // modulo 2^64; pq the exclusive per-word prefix of set bits)
    const u32 *pq = nullptr;
    const u64 *prow = nullptr;
    u32 *pout = nullptr;
    // Group mode (E_MASK, E_SCORES; 128 x 128 tiles): launch slot t is tile
    // (glist[t].x, glist[t].y) of the gathered operands, of group glist[t].z;
    // the epilogue sees the group's own rows, columns and output, and for
    // E_MASK its threshold gthr[2 g] (INT32_MAX: the group scores nothing).
    const uint4 *glist = nullptr;
    const GroupDesc *gdesc = nullptr;
    const int32_t *gthr = nullptr;
    u32 orgA = 0, orgB = 0;   // set in the kernel: the group's first A / B row
    // E_COUNTS (tryRadius counts without the score matrix; fp4 128 x 128
    // tiles): cpart[(J * nrad + k) * padA + i] = #{j of column tile J :
    // S(i, j) >= radii[k]} as a byte, from the count table of
    // launch_count_table (ctab / cneed / cpar).
    const unsigned char *ctab = nullptr, *cneed = nullptr;
    const int32_t *cpar = nullptr;
    u32 nrad = 0;
    unsigned char *cpart = nullptr;
};

// Largest span r_max - r_min + 1 of the radii the fused counts take (the
// count table lives in the tile's operand LDS after the k-loop).
constexpr u32 COUNT_SPAN_MAX = 20480;
constexpr u32 MAX_ROUND_RADII_FUSED = 64;

// First kw words of cell-major rows of stride ldw -> stage-major P (rows padded
// to a multiple of STAGE_PAD; kw a multiple of 32).  With rowidx, row r of P is
// row rowidx[r] of mat.
void launch_to_stage_major(const u32 *mat, u32 n, u32 ldw, u32 kw, u32 *P, hipStream_t s,
                           const u32 *rowidx = nullptr, u32 nsrc = 0xFFFFFFFFu);
bool mfma_enabled();
// Row popcounts of the first n rows of a stage-major matrix whose stages hold
// `stride` rows.
void launch_stage_popc(const u32 *P, u32 n, u32 stride, u32 ldw, u32 *pop, hipStream_t s);

// fp4 expansion of a stage-major copy (rows padded to pad, `stages` 128-site
// stages): every site becomes a nibble, 1.0 (0b0010) or 0; fp4_bytes() bytes.
inline size_t fp4_bytes(u32 pad, u32 stages) { return (size_t)stages * 2 * pad * 32; }
void launch_expand_fp4(const u32 *P, u32 pad, u32 stages, u32 *E, hipStream_t s);
void launch_mfma_tiles(int epi, const MfmaArgs &a, hipStream_t s);
// The same without timing anything: the autotuner's choice for the shape
// class when it has one, else the fp4 tiles (the int8 tiles when the fp4
// expansions EA / EB are not given or the sums could pass 2^24); creates no
// events and allocates nothing.
void launch_mfma_tiles_untuned(int epi, const MfmaArgs &a, hipStream_t s);
// Group-mode launch over `nslots` entries of a.glist (E_MASK or E_SCORES):
// the fp4 tiles when a.EA / a.EB are given and the sums stay exact, else the
// int8 tiles; allocates nothing and times nothing.
void launch_mfma_groups(int epi, const MfmaArgs &a, u32 nslots, hipStream_t s);
// tryRadius counts without the score matrix: the fp4 tiles (a.EA / a.EB given,
// sums exact, nradii <= 64) with the E_COUNTS epilogue into a.cpart
// (ceil(nB / 128) x nrad x padA bytes), then their sum over the column tiles
// into counts[k * R + q].  Allocates nothing.
bool fused_counts_ok(const MfmaArgs &a, u32 nradii);
// The count table of radii (ascending, nradii <= 64): ctab (COUNT_SPAN_MAX
// bytes), cneed (64 bytes), cpar (2 int32), all device memory.
void launch_count_table(const int32_t *radii, u32 nradii, unsigned char *ctab, unsigned char *cneed, int32_t *cpar,
                        hipStream_t s);
void launch_mfma_counts(const MfmaArgs &a, u32 *counts, u32 R, hipStream_t s);
// JSON report of the autotuner's choices (shape class -> variant, timings),
// followed by the notes other tuners added (one JSON object each).
std::string mfma_tuning_report();
void add_tuning_note(const std::string &json_object);

}  // namespace casphy
