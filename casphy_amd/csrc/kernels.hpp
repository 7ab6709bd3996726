// kernels.hpp -- layout conversion and list-compaction kernels (declarations).
#pragma once

#include "common.hpp"

namespace casphy {

// .bmat word-major [W][n] -> cell-major [n][ldw] with zero padding words.
void launch_transpose_wordmajor(const u32 *in, u32 W, u32 n, u32 *out, u32 ldw, hipStream_t s);
// cell-major [n][W] -> [n][ldw] with zero padding words.
void launch_pad_rows(const u32 *in, u32 W, u32 n, u32 *out, u32 ldw, hipStream_t s);

// DNA byte codes [n][nwRaw]: sets *flag (device u32, pre-zeroed) when some code
// has more than one nucleotide bit (IUPAC ambiguity) -> 2-bit form unusable.
void launch_dna_detect(const u32 *raw, u64 nwords, u32 *flag, hipStream_t s);
// 2-bit + mask form: per 32 positions three words (hi, lo, mask); rows of ldw.
void launch_dna_pack2(const u32 *raw, u32 n, u32 nwRaw, u32 *out, u32 ldw, hipStream_t s);
// nibble form: 8 positions per word (low nibble of each byte code); rows of ldw.
void launch_dna_pack_nib(const u32 *raw, u32 n, u32 nwRaw, u32 *out, u32 ldw, hipStream_t s);

// Radius mask -> CSR lists.
// skip (device, may be null): when it holds INT32_MAX every count is 0 and the
// mask is not read (a radius-mask launch that scored nothing).
void launch_row_popc(const u32 *mask, u32 n, u32 mstride, u32 words, u32 *counts, hipStream_t s,
                     const int32_t *skip = nullptr);
// Rows of several independent masks (the groups of a group-mode search): row
// r's mask starts at word moff[r] and has words[r] words; it counts 0 and is
// not read when gthr[2 group[r]] is INT32_MAX (its group scored nothing).
struct RowMap {
    const u64 *moff = nullptr;
    const u32 *words = nullptr;
    const u32 *group = nullptr;
    const int32_t *gthr = nullptr;
};
void launch_row_popc_map(const u32 *mask, u32 n, const RowMap &map, u32 *counts, hipStream_t s);
void launch_compact_map(const u32 *mask, u32 n, const RowMap &map, const u64 *offsets, u32 *idx, hipStream_t s,
                        u64 cap = ~0ull);
// Exclusive scan of n u32 counts into n+1 u64 offsets (offsets[n] = total).
void launch_scan_counts(const u32 *counts, u32 n, u64 *offsets, hipStream_t s);
// Ids at positions >= cap are not written (cap = capacity of idx).
void launch_compact(const u32 *mask, u32 n, u32 mstride, u32 words, const u64 *offsets, u32 *idx,
                    hipStream_t s, u64 cap = ~0ull);

// counts[k*R + q] = #{j < n : S[q*ldS + j] >= radii[k]} for k < nradii
// (<= MAX_ROUND_RADII; one counting pass per group of MAX_RADII_K radii).
constexpr int MAX_RADII_K = 16;
constexpr int MAX_ROUND_RADII = 64;
void launch_count_radii(const int32_t *S, u32 R, u32 n, u64 ldS, const int32_t *radii, u32 nradii, u32 *counts,
                        hipStream_t s);
// The same over rows of different lengths: row q is the rowlen[q] scores at
// S + rowoff[q].
void launch_count_radii_rows(const int32_t *S, u32 R, const u64 *rowoff, const u32 *rowlen, const int32_t *radii,
                             u32 nradii, u32 *counts, hipStream_t s);

// autoSplit choice (nradii <= MAX_ROUND_RADII): out[0] = radii[k] and out[1] = k for the first k with
// sum over q of counts[k*R + q] <= smax, or out[0] = INT32_MAX, out[1] = -1
// when none (or !any).
void launch_autosplit_choose(const u32 *counts, u32 R, const int32_t *radii, u32 nradii, u64 smax, int any,
                             int32_t *out, hipStream_t s);
// The choice of every group g at once: its queries are the columns
// qoff[g] .. qoff[g+1] of counts (R columns per radius), its limits smax[g] and
// any[g]; the result goes to out[2g], out[2g+1].
void launch_autosplit_choose_groups(const u32 *counts, u32 R, const u32 *qoff, u32 ngroups, const int32_t *radii,
                                    u32 nradii, const u64 *smax, const int32_t *any, int32_t *out, hipStream_t s);

// Scan form of tryRadius for up to SCAN_Q query rows (qrows: Q x ldw, cell-major):
// counts[k*R + q] = #{r < n : <mat_r, q-th query> >= radii[k]} for q < Q, k < nradii;
// S (Q x n, may be null) receives the scores; part holds scan_grid(n) x 64 u32 of
// per-workgroup counts.  For rows of 256..4,096 words.  ev0/ev1 (may be null)
// bracket the scoring kernel.
constexpr u32 SCAN_Q = 4;
bool scan_fits(u32 ldw);
u32 scan_grid(u32 n);
void launch_scan_bit(const u32 *mat, u32 n, u32 ldw, const u32 *qrows, u32 Q, const int32_t *radii, u32 nradii,
                     u32 *counts, u32 R, int32_t *S, u32 *part, hipStream_t s, hipEvent_t ev0 = nullptr,
                     hipEvent_t ev1 = nullptr);

// Word copies in a kernel (either side may be mapped pinned host memory).
void launch_copy_words(const u32 *src, u32 *dst, u32 n, hipStream_t s);
// out[0..1] = choice[0..1], out[2..3] = *total (low, high), out[4..] = counts.
void launch_publish(const int32_t *choice, const u64 *total, const u32 *counts, u32 ncounts, u32 *out,
                    hipStream_t s);

// size[r] = min(p, ntip - p) with p = popcount of row r (u32 arithmetic).
void launch_row_size(const u32 *mat, u32 n, u32 ldw, u32 ntip, u32 *size, hipStream_t s);

void launch_fill_u32(u32 *p, u64 count, u32 value, hipStream_t s);
void launch_fill_u64(u64 *p, u64 count, u64 value, hipStream_t s);

}  // namespace casphy
