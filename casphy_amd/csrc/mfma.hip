// mfma.hip -- exact AND-count on the int8 matrix cores (gfx950).
//
// popcount(a & b) over bit-packed sites equals the dot product of the 0/1
// byte vectors, and v_mfma_i32_32x32x32_i8 accumulates such dot products in
// int32 exactly (at most 2^31 / 1 = every realistic site count).  The kernel:
//   * reads the packed bits in a stage-major layout P[stage][row][4 u32]
//     (128 sites = 16 bytes per row per stage, 2 KB contiguous per 128 rows),
//     so every global load is a fully coalesced 8-byte load;
//   * expands each 32-bit word into 32 bytes of 0/1 in registers
//     ((nibble * 0x00204081) & 0x01010101 per output word) and stores them in
//     LDS rows of 128 + 16 bytes (the pad spreads rows over the LDS banks);
//   * 4 waves per workgroup, each owning a 64 x 64 block of the 128 x 128 pair
//     tile as 2 x 2 MFMA 32x32 blocks (64 int32 accumulators per lane);
//   * fuses the same epilogues as the VALU tiles: the radius bitmask words are
//     assembled with one 64-lane ballot per accumulator register (the two
//     halves of the wave hold rows m and m+4 of one 32-column word), gathered
//     into 32 lanes with writelane and written by one store per block; the
//     mirrored words come from the per-lane column bits and one cross-half
//     shuffle.
// Accumulator layout of a 32x32 MFMA block: lane l, register v holds row
// 8*(v/4) + 4*(l/32) + v%4 and column l%32.  The A and B operands use the same
// (lane, byte) -> k mapping, so the product does not depend on how the
// hardware orders k inside an instruction.
#include "mfma.hpp"

#include <atomic>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

namespace casphy {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));

// (j, i) with i <= j of the t-th entry of the upper triangle in column order.
__device__ __forceinline__ void tri_index(u32 t, u32 &i, u32 &j)
{
    u32 c = (u32)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((u64)c * (c + 1) / 2 > t) --c;
    while ((u64)(c + 1) * (c + 2) / 2 <= t) ++c;
    j = c;
    i = t - (u32)((u64)c * (c + 1) / 2);
}

// Tile (I, J) of launch slot t.  With a.sup = S > 0 the tiles are numbered
// by S x S super-tiles (upper triangle of super-tiles for self scoring), so
// the workgroups one XCD holds at a time -- a contiguous run of slots -- share
// about 2 S row blocks instead of sweeping whole columns; slots of a super-tile
// that fall outside the triangle or the matrix hold no tile (false).
__device__ __forceinline__ bool tile_of(const MfmaArgs &a, u32 t, u32 &I, u32 &J)
{
    const u32 S = a.sup;
    if (S == 0) {
        if (a.triangle) {
            tri_index(t, I, J);
        } else {
            I = t / a.ntB;
            J = t - I * a.ntB;
        }
        return true;
    }
    const u32 S2 = S * S, u = t / S2, l = t - u * S2;
    u32 SI, SJ;
    if (a.triangle) {
        tri_index(u, SI, SJ);
    } else {
        const u32 nSB = (a.ntB + S - 1) / S;
        SI = u / nSB;
        SJ = u - SI * nSB;
    }
    I = SI * S + l / S;
    J = SJ * S + l % S;
    return a.triangle ? (I <= J && J < a.ntA) : (I < a.ntA && J < a.ntB);
}


// A radius-mask launch whose device threshold is INT32_MAX (an autoSplit round
// in which no radius qualified) scores nothing: no pair can reach it, and the
// counts and compaction that follow skip the rows without reading the mask.
template <int EPI>
__device__ __forceinline__ bool mask_skipped(const MfmaArgs &a)
{
    return EPI == E_MASK && a.thr_dev && *a.thr_dev == 0x7FFFFFFF;
}

constexpr int NTHR = 256;

// Tile (I, J) of launch slot t and, in group mode, the group's view for the
// epilogue: its row counts, output and threshold, and the origin that turns
// rows of the gathered operands into the group's own rows.  False: the slot
// holds no tile, or the tile's group scores nothing.
template <int EPI>
__device__ __forceinline__ bool resolve_tile(MfmaArgs &a, u32 t, u32 &I, u32 &J)
{
    if (!a.glist) return !mask_skipped<EPI>(a) && tile_of(a, t, I, J);
    const uint4 d = a.glist[t];
    const GroupDesc g = a.gdesc[d.z];
    I = d.x;
    J = d.y;
    a.orgA = g.a0;
    a.orgB = g.b0;
    a.nA = g.nA;
    a.nB = g.nB;
    if (EPI == E_MASK) {
        a.mask += g.off;
        a.mstride = g.ld;
        if (a.gthr) a.thr = a.gthr[2 * d.z];
        a.thr_dev = nullptr;
        return a.thr != 0x7FFFFFFF;
    }
    a.S += g.off;
    a.ldS = g.ld;
    return true;
}

__global__ void k_to_stage_major(const u32 *mat, const u32 *rowidx, u32 n, u32 ldw, u32 kw, u32 nrows_pad, u32 *P,
                                 u32 nsrc)
{
    // P[s][r][q] = mat[r][4s + q] for the first kw words of rows of stride ldw;
    // rows >= n, and gathered rows naming a row >= nsrc, are zero
    const u32 stages = kw / 4;
    const u64 total = (u64)stages * nrows_pad;
    for (u64 x = blockIdx.x * (u64)blockDim.x + threadIdx.x; x < total; x += (u64)gridDim.x * blockDim.x) {
        const u32 s = (u32)(x / nrows_pad), r = (u32)(x - (u64)s * nrows_pad);
        uint4 v = make_uint4(0, 0, 0, 0);
        u32 src = r < n ? (rowidx ? rowidx[r] : r) : 0xFFFFFFFFu;
        if (src >= nsrc) src = 0xFFFFFFFFu;
        if (src != 0xFFFFFFFFu) v = *reinterpret_cast<const uint4 *>(mat + (size_t)src * ldw + 4 * s);
        reinterpret_cast<uint4 *>(P)[x] = v;
    }
}

__device__ __forceinline__ u32 spread4(u32 nib) { return (nib * 0x00204081u) & 0x01010101u; }

__device__ __forceinline__ uint4 spread16(u32 w)  // bits 0..15 -> 16 bytes
{
    return make_uint4(spread4(w & 15), spread4((w >> 4) & 15), spread4((w >> 8) & 15), spread4((w >> 12) & 15));
}

// Per-thread slice of one stage: one row of one operand (threads 0..127 the A
// rows, 128..255 the B rows), KB sites of packed bits.  Consecutive lanes own
// consecutive rows, so the global load is contiguous and the expanded bytes
// go to LDS rows LR bytes apart (LR = KB + 16), which keeps the 16-byte
// stores of 8 consecutive lanes on distinct banks.
template <int KB> struct Slice;
template <> struct Slice<128> {
    uint4 v;
    __device__ __forceinline__ void load(const u32 *P, size_t pad, u32 tile, u32 s, u32 row)
    {
        v = reinterpret_cast<const uint4 *>(P)[(size_t)s * pad + (size_t)tile * TILE + row];
    }
    __device__ __forceinline__ void store(unsigned char *dst, u32 row) const
    {
        uint4 *d = reinterpret_cast<uint4 *>(dst + row * (128 + 16));
        d[0] = spread16(v.x);
        d[1] = spread16(v.x >> 16);
        d[2] = spread16(v.y);
        d[3] = spread16(v.y >> 16);
        d[4] = spread16(v.z);
        d[5] = spread16(v.z >> 16);
        d[6] = spread16(v.w);
        d[7] = spread16(v.w >> 16);
    }
};
template <> struct Slice<64> {
    uint2 v;
    __device__ __forceinline__ void load(const u32 *P, size_t pad, u32 tile, u32 s, u32 row)
    {
        // 64-site stage s = half (s & 1) of 128-site stage s >> 1
        v = reinterpret_cast<const uint2 *>(P)[((size_t)(s >> 1) * pad + (size_t)tile * TILE + row) * 2 + (s & 1)];
    }
    __device__ __forceinline__ void store(unsigned char *dst, u32 row) const
    {
        uint4 *d = reinterpret_cast<uint4 *>(dst + row * (64 + 16));
        d[0] = spread16(v.x);
        d[1] = spread16(v.x >> 16);
        d[2] = spread16(v.y);
        d[3] = spread16(v.y >> 16);
    }
};

// Radius-mask compares of accumulator registers V..15 of a 32 x 32 block.
// Register V holds row t = 8 (V / 4) + V % 4 in the low half of the wave and
// row t + 4 in the high half, one column per lane: the ballot of the compare
// is the mask word of both rows, which v_writelane_b32 (lane index an
// immediate, no memory access) puts into lanes t and t + 4 of rowword.
// colbits collects the lane's own hits (bit t) for the mirrored words.
template <int V>
__device__ __forceinline__ void mask_rows(const v16i &acc, int32_t thr, unsigned long long jmask, u32 &rowword,
                                          u32 &colbits)
{
    constexpr int t = 8 * (V / 4) + (V % 4);
    const bool hit = acc[V] >= thr;
    const unsigned long long b = __ballot(hit) & jmask;
    asm("v_writelane_b32 %0, %1, %2" : "+v"(rowword) : "s"((u32)b), "i"(t));
    asm("v_writelane_b32 %0, %1, %2" : "+v"(rowword) : "s"((u32)(b >> 32)), "i"(t + 4));
    colbits |= hit ? 1u << t : 0u;
    if constexpr (V + 1 < 16) mask_rows<V + 1>(acc, thr, jmask, rowword, colbits);
}

// Radius mask of one 32 x 32 block from fp32 accumulators (the fp4 tiles):
// every sum is an exact integer below 2^24, so `sum >= (float)thr` decides as
// the integer compare does (a threshold beyond +-2^24 rounds to a value that
// still lies beyond every sum).  Register V holds rows t and t + 4 of the block
// (t = 8 (V / 4) + V % 4) in the two half-waves: one ballot is both rows' words;
// they gather into lanes t and t + 4 of two registers (two chains of
// writelanes, so no writelane waits on the one before it).  The writelanes
// read the ballot through the scalar AND with the column mask: on this target
// a writelane that reads a scalar register straight after a vector compare
// wrote it reads the old value, so the scalar instruction between them is
// part of the contract (a compiler that dropped the AND would break it; the
// column mask is a runtime ballot, so it cannot).  Each lane's own compares
// (its column of the block, for the mirrored words) collect in colbits by one
// add with carry per register: colbits = 2 colbits + compare, registers 15
// down to 0, so bit V of colbits is register V's compare.
template <int V>
__device__ __forceinline__ void mask_rows_f(const v16f &acc, float thr, unsigned long long jmask, u32 &rw0, u32 &rw1,
                                            u32 &colbits)
{
    constexpr int t = 8 * (V / 4) + (V % 4);
    const unsigned long long hit = __ballot(acc[V] >= thr);
    const unsigned long long b = hit & jmask;
    asm("v_writelane_b32 %0, %1, %2" : "+v"(rw0) : "s"((u32)b), "i"(t));
    asm("v_writelane_b32 %0, %1, %2" : "+v"(rw1) : "s"((u32)(b >> 32)), "i"(t + 4));
    unsigned long long carry;
    asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(colbits), "=s"(carry) : "v"(colbits), "s"(hit));
    if constexpr (V > 0) mask_rows_f<V - 1>(acc, thr, jmask, rw0, rw1, colbits);
}

// Epilogue of one 32 x 32 block whose rows start at r0 and columns at c0
// (mirror: the tile is off the diagonal of a triangle, so the transposed
// pairs are written too).  Row blocks of the transfer-distance batches are
// 128 rows (tile_rep per 128-row block).
template <int EPI>
__device__ __forceinline__ void block_epilogue(const MfmaArgs &a, const v16i &acc, u32 r0, u32 c0, u32 lane,
                                               bool mirror, int32_t thr)
{
    const u32 half = lane >> 5, col = lane & 31, j = c0 + col;
    if (EPI == E_MASK) {
        // the 32 row words gather in lanes 0..31 and leave in one store; the
        // mirrored words (one per column) join each lane's 16 compares with the
        // other half's 16
        const unsigned long long jmask = __ballot(j < a.nB);
        u32 rowword = 0, colbits = 0;
        mask_rows<0>(acc, thr, jmask, rowword, colbits);
        const u32 wc = c0 / 32;
        if (lane < 32 && r0 + lane < a.nA && wc < a.mstride) a.mask[(size_t)(r0 + lane) * a.mstride + wc] = rowword;
        if (mirror) {
            colbits <<= 4 * half;
            const u32 rows = r0 < a.nA ? a.nA - r0 : 0u;  // rows past nA score the zero padding
            colbits &= rows >= 32 ? 0xFFFFFFFFu : (1u << rows) - 1u;
            colbits |= __shfl_xor(colbits, 32);
            if (half == 0 && j < a.nB && r0 / 32 < a.mstride) a.mask[(size_t)j * a.mstride + r0 / 32] = colbits;
        }
    } else if (EPI == E_SCORES) {
        if (r0 + 32 <= a.nA && c0 + 32 <= a.nB) {
            // whole block in range: no per-element tests, row offsets fixed per register
            const size_t ld = a.ldS;
            int32_t *p = a.S + (size_t)(r0 + 4 * half) * ld + j;
#pragma unroll
            for (int v = 0; v < 16; ++v) p[(size_t)(8 * (v / 4) + (v % 4)) * ld] = acc[v];
            if (mirror) {
                // registers 4g .. 4g+3 are 4 consecutive rows: one 16-byte store
                // per group into row j when the rows of S are 16-byte aligned
                int32_t *pm = a.S + (size_t)j * ld + r0 + 4 * half;
                if (((ld & 3) | (reinterpret_cast<uintptr_t>(a.S) & 15)) == 0) {
#pragma unroll
                    for (int g = 0; g < 4; ++g)
                        *reinterpret_cast<v4i *>(pm + 8 * g) = v4i{acc[4 * g], acc[4 * g + 1], acc[4 * g + 2],
                                                                    acc[4 * g + 3]};
                } else {
#pragma unroll
                    for (int v = 0; v < 16; ++v) pm[8 * (v / 4) + (v % 4)] = acc[v];
                }
            }
        } else {
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const u32 i = r0 + 8 * (v / 4) + 4 * half + (v % 4);
                if (i < a.nA && j < a.nB) {
                    a.S[(size_t)i * a.ldS + j] = acc[v];
                    if (mirror) a.S[(size_t)j * a.ldS + i] = acc[v];
                }
            }
        }
    } else if (EPI == E_LOCALMAX) {
        // per row: max over j > i of (S, smallest j), reduced over the 32 lanes of a half
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const u32 i = r0 + 8 * (v / 4) + 4 * half + (v % 4);
            unsigned long long key = 0;
            if (j < a.nB && j > i) key = ((unsigned long long)(u32)acc[v] << 32) | (0xFFFFFFFFu - j);
#pragma unroll
            for (int o = 1; o < 32; o <<= 1) {
                const unsigned long long q = __shfl_xor(key, o);
                key = q > key ? q : key;
            }
            if (col == 0 && i < a.nA && key) atomicMax((unsigned long long *)&a.keys[i], key);
        }
    } else if (EPI == E_ARGMAX) {
        // per column: max over the A rows i of (S, smallest i); a lane holds 16
        // rows of its column, the two halves meet in one exchange
        unsigned long long key = 0;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const u32 i = r0 + 8 * (v / 4) + 4 * half + (v % 4);
            const unsigned long long k = ((unsigned long long)(u32)acc[v] << 32) | (0xFFFFFFFFu - i);
            key = (i < a.nA && k > key) ? k : key;
        }
        const unsigned long long q = __shfl_xor(key, 32);
        key = q > key ? q : key;
        if (half == 0 && j < a.nB && key) atomicMax((unsigned long long *)&a.keys[j], key);
    } else if (EPI == E_PAIRS) {
        // shared-neighbour counts of the linked pairs i < j, at their place in
        // the upper part of row i's list; the 32 columns of a block are one
        // word of each row of the neighbour bits
        const u32 wc = c0 / 32, below = (1u << col) - 1u;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const u32 i = r0 + 8 * (v / 4) + 4 * half + (v % 4);
            if (i < a.nA && j < a.nB && i < j) {
                const size_t at = (size_t)i * a.mstride + wc;
                const u32 word = a.mask[at];
                if ((word >> col) & 1u)
                    a.pout[a.prow[i] + a.pq[at] + (u32)__builtin_popcount(word & below)] = (u32)acc[v];
            }
        }
    } else if (EPI == E_TBE && r0 < a.nA) {  // a block past the last row has nothing to write
        const u32 pb = j < a.nB ? a.popB[j] : 0u;
        const u32 bbase = a.tile_rep ? a.tile_rep[r0 >> 7] * a.bstride : 0u;
        u32 cmin = 0xFFFFFFFFu;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const u32 i = r0 + 8 * (v / 4) + 4 * half + (v % 4);
            const u32 out = i < a.nA ? (a.arow_map ? a.arow_map[i] : i) : 0xFFFFFFFFu;
            const bool valid = out != 0xFFFFFFFFu && j < a.nB;
            const u32 h = (valid ? a.popA[i] : 0u) + pb - 2u * (u32)acc[v];
            const u32 hc = a.ntip - h;
            u32 d = valid ? (h < hc ? h : hc) : 0xFFFFFFFFu;
            cmin = d < cmin ? d : cmin;
#pragma unroll
            for (int o = 1; o < 32; o <<= 1) {
                const u32 q = __shfl_xor(d, o);
                d = q < d ? q : d;
            }
            if (col == 0 && out != 0xFFFFFFFFu) atomicMin(&a.amin[out], d);
        }
        const u32 q = __shfl_xor(cmin, 32);
        cmin = q < cmin ? q : cmin;
        if (half == 0 && j < a.nB) atomicMin(&a.bmin[bbase + j], cmin);
    }
}

// Epilogue of one 32 x 32 block of the fp4 tiles (fp32 accumulators): the
// radius mask compares the sums as floats; every other epilogue takes them as
// the integers they are.
template <int EPI>
__device__ __forceinline__ void fp4_epilogue(const MfmaArgs &a, const v16f &acc, u32 r0, u32 c0, u32 lane, bool mirror,
                                             int32_t thr)
{
    if constexpr (EPI == E_MASK) {
        const u32 half = lane >> 5, j = c0 + (lane & 31);
        const float fthr = (float)thr;
        u32 rw0 = 0, rw1 = 0, colbits = 0;
        mask_rows_f<15>(acc, fthr, __ballot(j < a.nB), rw0, rw1, colbits);
        const u32 rowword = rw0 | rw1, wc = c0 / 32;
        if (lane < 32 && r0 + lane < a.nA && wc < a.mstride) a.mask[(size_t)(r0 + lane) * a.mstride + wc] = rowword;
        if (mirror) {
            // bit V -> row 8 (V / 4) + V % 4: nibble k of the 16 bits to byte k
            colbits = (colbits | (colbits << 8)) & 0x00FF00FFu;
            colbits = (colbits | (colbits << 4)) & 0x0F0F0F0Fu;
            colbits <<= 4 * half;
            const u32 rows = r0 < a.nA ? a.nA - r0 : 0u;  // rows past nA score the zero padding
            colbits &= rows >= 32 ? 0xFFFFFFFFu : (1u << rows) - 1u;
            colbits |= __shfl_xor(colbits, 32);
            if (half == 0 && j < a.nB && r0 / 32 < a.mstride) a.mask[(size_t)j * a.mstride + r0 / 32] = colbits;
        }
    } else {
        v16i ai;
#pragma unroll
        for (int v = 0; v < 16; ++v) ai[v] = (int)acc[v];
        block_epilogue<EPI>(a, ai, r0, c0, lane, mirror, thr);
    }
}

// KB sites per LDS stage, PF stages of global prefetch in registers, OCC
// workgroups per CU the launch bounds ask for.
template <int EPI, int KB, int PF, int OCC>
__global__ __launch_bounds__(NTHR, OCC) void k_mfma_tiles(MfmaArgs a)
{
    constexpr int LR = KB + 16;  // LDS bytes per row (the pad spreads rows over the banks)
    __shared__ __attribute__((aligned(16))) unsigned char As[2][TILE * LR];
    __shared__ __attribute__((aligned(16))) unsigned char Bs[2][TILE * LR];

    const u32 per = (a.ntiles + 7) / 8;
    const u32 t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (t >= a.ntiles) return;
    u32 I, J;
    if (!resolve_tile<EPI>(a, t, I, J)) return;
    const u32 tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const u32 wm = wave >> 1, wn = wave & 1;

    v16i acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[x][y][v] = 0;

    const u32 nst = a.stages * (128 / KB);  // a multiple of 8: rows are padded to 1,024 sites
    const bool opB = tid >= TILE;  // wave-uniform: waves 0-1 stage A rows, waves 2-3 B rows
    const u32 row = tid & (TILE - 1);
    const u32 *P = opB ? a.PB : a.PA;
    const size_t pad = opB ? a.padB : a.padA;
    const u32 tile = opB ? J : I;
    unsigned char(*Ls)[TILE * LR] = opB ? Bs : As;
    Slice<KB> g[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) g[u].load(P, pad, tile, u, row);
    g[0].store(Ls[0], row);
    if (PF < (int)nst) g[0].load(P, pad, tile, PF, row);
    __syncthreads();
    const u32 arow = wm * 64 + (lane & 31), brow = wn * 64 + (lane & 31), koff = 16 * (lane >> 5);
    for (u32 s0 = 0; s0 < nst; s0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const u32 s = s0 + u;
            const int buf = s & 1;
#pragma unroll
            for (int kk = 0; kk < KB; kk += 32) {
                v4i fa[2], fb[2];
#pragma unroll
                for (int x = 0; x < 2; ++x) {
                    fa[x] = *reinterpret_cast<const v4i *>(&As[buf][(arow + 32 * x) * LR + kk + koff]);
                    fb[x] = *reinterpret_cast<const v4i *>(&Bs[buf][(brow + 32 * x) * LR + kk + koff]);
                }
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y)
                        acc[x][y] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[x], fb[y], acc[x][y], 0, 0, 0);
            }
            const int nx = (u + 1) % PF;  // static after unrolling
            if (s + 1 < nst) {
                g[nx].store(Ls[buf ^ 1], row);
                if (s + 1 + PF < nst) g[nx].load(P, pad, tile, s + 1 + PF, row);
            }
            __syncthreads();
        }
    }

    const int32_t thr = EPI == E_MASK && a.thr_dev ? *a.thr_dev : a.thr;
    const bool mirror = a.triangle && I != J;
    const u32 er0 = I * TILE - a.orgA + wm * 64, ec0 = J * TILE - a.orgB + wn * 64;  // the group's own rows
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
            block_epilogue<EPI>(a, acc[x][y], er0 + 32 * x, ec0 + 32 * y, lane, mirror, thr);
}

// ---------------------------------------------------------------------------
// Wide LDS variant: a workgroup of 4 waves (2 x 2) owns a 256 x 256 pair tile
// and each wave 128 x 128 pairs as 4 x 4 MFMA blocks, so every 16-byte operand
// read from LDS feeds 4 MFMAs (twice the 2 x 2 blocking): per 64-site stage
// the CU moves 32 KB of expanded bytes into LDS and 64 KB out of it for 128
// MFMAs.  One workgroup per CU (256 accumulators per lane); the next stage's
// packed bits are prefetched into registers while the current one is used.
constexpr u32 WT = 256;  // workgroup tile

template <int EPI>
__global__ __launch_bounds__(256, 1) void k_mfma_wide(MfmaArgs a)
{
    constexpr int KB = 64, LR = KB + 16;  // sites per LDS stage; LDS bytes per row
    __shared__ __attribute__((aligned(16))) unsigned char As[2][WT * LR];
    __shared__ __attribute__((aligned(16))) unsigned char Bs[2][WT * LR];

    const u32 per = (a.ntiles + 7) / 8;
    const u32 t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (t >= a.ntiles || mask_skipped<EPI>(a)) return;
    u32 I, J;
    if (!tile_of(a, t, I, J)) return;
    const u32 tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const u32 wm = wave >> 1, wn = wave & 1;

    v16i acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[x][y][v] = 0;

    // thread tid stages row tid of the A tile and of the B tile: 64 sites =
    // half (s & 1) of 128-site stage s >> 1, 8 bytes each
    const u32 nst = a.stages * 2;
    const uint2 *PA = reinterpret_cast<const uint2 *>(a.PA) + ((size_t)I * WT + tid) * 2;
    const uint2 *PB = reinterpret_cast<const uint2 *>(a.PB) + ((size_t)J * WT + tid) * 2;
    auto ld = [&](const uint2 *P, size_t pad, u32 s) { return P[(size_t)(s >> 1) * pad * 2 + (s & 1)]; };
    auto st = [&](unsigned char *dst, uint2 v) {
        uint4 *d = reinterpret_cast<uint4 *>(dst + tid * LR);
        d[0] = spread16(v.x);
        d[1] = spread16(v.x >> 16);
        d[2] = spread16(v.y);
        d[3] = spread16(v.y >> 16);
    };
    uint2 ga = ld(PA, a.padA, 0), gb = ld(PB, a.padB, 0);
    st(As[0], ga);
    st(Bs[0], gb);
    if (nst > 1) {
        ga = ld(PA, a.padA, 1);
        gb = ld(PB, a.padB, 1);
    }
    __syncthreads();
    const u32 arow = wm * 128 + (lane & 31), brow = wn * 128 + (lane & 31), koff = 16 * (lane >> 5);
    for (u32 s = 0; s < nst; ++s) {
        const int buf = s & 1;
#pragma unroll
        for (int kk = 0; kk < KB; kk += 32) {
            v4i fa[4], fb[4];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                fa[x] = *reinterpret_cast<const v4i *>(&As[buf][(arow + 32 * x) * LR + kk + koff]);
                fb[x] = *reinterpret_cast<const v4i *>(&Bs[buf][(brow + 32 * x) * LR + kk + koff]);
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[x], fb[y], acc[x][y], 0, 0, 0);
        }
        if (s + 1 < nst) {
            st(As[buf ^ 1], ga);
            st(Bs[buf ^ 1], gb);
            if (s + 2 < nst) {
                ga = ld(PA, a.padA, s + 2);
                gb = ld(PB, a.padB, s + 2);
            }
        }
        __syncthreads();
    }

    const int32_t thr = EPI == E_MASK && a.thr_dev ? *a.thr_dev : a.thr;
    const bool mirror = a.triangle && I != J;
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
            block_epilogue<EPI>(a, acc[x][y], I * WT + wm * 128 + 32 * x, J * WT + wn * 128 + 32 * y, lane,
                                     mirror, thr);
}

// Wide tile with two waves per SIMD: 8 waves (2 x 4) per 256 x 256 tile, each
// 128 x 64 pairs as 4 x 2 MFMA blocks (128 accumulators per lane).  Per
// 64-site stage the CU moves 32 KB into LDS and 96 KB out of it for 128 MFMAs;
// threads 0..255 stage A rows and 256..511 B rows, one row each.
template <int EPI>
__global__ __launch_bounds__(512, 1) void k_mfma_wide8(MfmaArgs a)
{
    constexpr int KB = 64, LR = KB + 16;
    __shared__ __attribute__((aligned(16))) unsigned char As[2][WT * LR];
    __shared__ __attribute__((aligned(16))) unsigned char Bs[2][WT * LR];

    const u32 per = (a.ntiles + 7) / 8;
    const u32 t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (t >= a.ntiles || mask_skipped<EPI>(a)) return;
    u32 I, J;
    if (!tile_of(a, t, I, J)) return;
    const u32 tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const u32 wm = wave >> 2, wn = wave & 3;

    v16i acc[4][2];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[x][y][v] = 0;

    const u32 nst = a.stages * 2;
    const bool opB = tid >= WT;  // wave-uniform
    const u32 row = tid & (WT - 1);
    const uint2 *P = reinterpret_cast<const uint2 *>(opB ? a.PB : a.PA) + ((size_t)(opB ? J : I) * WT + row) * 2;
    const size_t pad2 = (size_t)(opB ? a.padB : a.padA) * 2;
    unsigned char(*Ls)[WT * LR] = opB ? Bs : As;
    auto ld = [&](u32 s) { return P[(size_t)(s >> 1) * pad2 + (s & 1)]; };
    auto st = [&](unsigned char *dst, uint2 v) {
        uint4 *d = reinterpret_cast<uint4 *>(dst + row * LR);
        d[0] = spread16(v.x);
        d[1] = spread16(v.x >> 16);
        d[2] = spread16(v.y);
        d[3] = spread16(v.y >> 16);
    };
    uint2 g0 = ld(0), g1 = nst > 1 ? ld(1) : make_uint2(0, 0);
    st(Ls[0], g0);
    if (nst > 2) g0 = ld(2);
    __syncthreads();
    const u32 arow = wm * 128 + (lane & 31), brow = wn * 64 + (lane & 31), koff = 16 * (lane >> 5);
    for (u32 s = 0; s < nst; ++s) {
        const int buf = s & 1;
#pragma unroll
        for (int kk = 0; kk < KB; kk += 32) {
            v4i fa[4], fb[2];
#pragma unroll
            for (int x = 0; x < 4; ++x)
                fa[x] = *reinterpret_cast<const v4i *>(&As[buf][(arow + 32 * x) * LR + kk + koff]);
#pragma unroll
            for (int y = 0; y < 2; ++y)
                fb[y] = *reinterpret_cast<const v4i *>(&Bs[buf][(brow + 32 * y) * LR + kk + koff]);
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[x], fb[y], acc[x][y], 0, 0, 0);
        }
        if (s + 1 < nst) {
            // stage s + 1 was loaded two iterations ago (g1 for odd s + 1, g0 for even)
            st(Ls[buf ^ 1], (s & 1) ? g0 : g1);
            if (s + 3 < nst) {
                if (s & 1) g0 = ld(s + 3);
                else g1 = ld(s + 3);
            }
        }
        __syncthreads();
    }

    const int32_t thr = EPI == E_MASK && a.thr_dev ? *a.thr_dev : a.thr;
    const bool mirror = a.triangle && I != J;
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
            block_epilogue<EPI>(a, acc[x][y], I * WT + wm * 128 + 32 * x, J * WT + wn * 64 + 32 * y, lane,
                                     mirror, thr);
}

// ---------------------------------------------------------------------------
// fp4 variant.  A 0/1 site becomes the fp4 (e2m1) value 1.0 = 0b0010 or 0.0,
// and v_mfma_scale_f32_32x32x64_f8f6f4 with both operands in fp4 and unit
// scales (exponent 127) adds 64 products of 1.0 x 1.0 per instruction to an
// fp32 accumulator: every partial sum is an integer below 2^24, so the fp32
// sums are exact and equal the AND counts.  The fp4 rate is twice the int8
// rate, and an operand carries 64 sites in 16 bytes per lane instead of 32.
// The expansion to nibbles runs once per launch into HBM (stage-major, 64-site
// chunks of 32 bytes per row), so the tile loop only copies bytes to LDS.
typedef int v8i __attribute__((ext_vector_type(8)));

__device__ __forceinline__ u32 nib8(u32 b)  // bits 0..7 -> nibbles 0..7 of value 2 or 0
{
    u32 x = b & 0xFFu;
    x = (x | (x << 12)) & 0x000F000Fu;
    x = (x | (x << 6)) & 0x03030303u;
    x = (x | (x << 3)) & 0x11111111u;
    return x << 1;
}

// P[s][r] (128 sites, 16 bytes) -> E chunk 2s (sites 0..63) and 2s + 1 (sites
// 64..127); chunk c of row r is the 32 bytes at E + (c * pad + r) * 32.
__global__ void k_expand_fp4(const u32 *P, u32 pad, u32 stages, u32 *E)
{
    const u64 total = (u64)stages * pad;
    for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < total; t += (u64)gridDim.x * blockDim.x) {
        const u32 s = (u32)(t / pad), r = (u32)(t - (u64)s * pad);
        const uint4 v = reinterpret_cast<const uint4 *>(P)[t];
        uint4 *e0 = reinterpret_cast<uint4 *>(E) + ((size_t)(2 * s) * pad + r) * 2;
        uint4 *e1 = reinterpret_cast<uint4 *>(E) + ((size_t)(2 * s + 1) * pad + r) * 2;
        e0[0] = make_uint4(nib8(v.x), nib8(v.x >> 8), nib8(v.x >> 16), nib8(v.x >> 24));
        e0[1] = make_uint4(nib8(v.y), nib8(v.y >> 8), nib8(v.y >> 16), nib8(v.y >> 24));
        e1[0] = make_uint4(nib8(v.z), nib8(v.z >> 8), nib8(v.z >> 16), nib8(v.z >> 24));
        e1[1] = make_uint4(nib8(v.w), nib8(v.w >> 8), nib8(v.w >> 16), nib8(v.w >> 24));
    }
}

// tryRadius counts of one 128 x 128 tile without storing its scores: each
// score's bin idx = #{k : radii[k] <= score} comes from the count table
// (copied into the B operand's LDS), every row's bins are counted in the A
// operand's LDS (two 16-bit bins per word; a bin holds at most 128), then one
// thread per row turns its bins into suffix sums and writes, per radius k, the
// row's count of columns with idx >= cneed[k] as one byte of cpart.  The
// operand LDS is free: every wave has left the k-loop.
__device__ __forceinline__ void count_epilogue(const MfmaArgs &a, const v16f (&acc)[2][2], u32 I, u32 J, u32 wm,
                                               u32 wn, u32 lane, u32 tid, u32 *H, unsigned char *T)
{
    constexpr u32 HW = 33;  // words per row: bins 0 .. 65 as 16-bit halves
    const int32_t cmin = a.cpar[0];
    const u32 cspan = min((u32)a.cpar[1], COUNT_SPAN_MAX), nrad = a.nrad;
    __syncthreads();
    for (u32 w = tid; w < TILE * HW; w += NTHR) H[w] = 0u;
    for (u32 x = tid; x < cspan; x += NTHR) T[x] = a.ctab[x];
    __syncthreads();
    const u32 half = lane >> 5, col = lane & 31;
    const u32 topw = nrad >> 1, tops = 16 * (nrad & 1);
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const bool jv = J * TILE + wn * 64 + 32 * y + col < a.nB;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const u32 rl0 = wm * 64 + 32 * x + 8 * (v / 4) + (v % 4), rl = rl0 + 4 * half;
                const int32_t d = (int32_t)acc[x][y][v] - cmin;
                const u32 idx = d < 0 ? 0u : (u32)d >= cspan ? nrad : (u32)T[d];
                const bool valid = jv && idx && I * TILE + rl < a.nA;
                // scores above every radius (the common case) are counted per
                // half-wave row with one ballot instead of 32 conflicting atomics
                const unsigned long long top = __ballot(valid && idx == nrad);
                if (lane == 0 && top) {
                    const u32 lo = (u32)__builtin_popcountll(top & 0xFFFFFFFFull), hi = (u32)__builtin_popcountll(top >> 32);
                    if (lo) atomicAdd(&H[rl0 * HW + topw], lo << tops);
                    if (hi) atomicAdd(&H[(rl0 + 4) * HW + topw], hi << tops);
                }
                if (valid && idx != nrad) atomicAdd(&H[rl * HW + (idx >> 1)], 1u << (16 * (idx & 1)));
            }
        }
    __syncthreads();
    if (tid < TILE) {
        // the row's bins in registers, the suffix sum from the top bin down; the
        // radii (cneed non-increasing from the last one) are uniform over the row
        // threads, so the inner loop does not diverge
        u32 bins[HW];
#pragma unroll
        for (int w = 0; w < (int)HW; ++w) bins[w] = H[tid * HW + w];
        unsigned char *out = a.cpart + (size_t)J * nrad * a.padA + I * TILE + tid;
        u32 run = 0;
        int k = (int)nrad - 1;
#pragma unroll
        for (int b = 64; b >= 1; --b) {
            run += (bins[b >> 1] >> (16 * (b & 1))) & 0xFFFFu;
            while (k >= 0 && a.cneed[k] == (u32)b) {
                out[(size_t)k * a.padA] = (unsigned char)run;
                --k;
            }
        }
    }
}

// 128 x 128 pair tile, 4 waves of 2 x 2 blocks, LDS stages of 128 sites (two
// 64-site chunks, rows of 64 + 16 bytes), one stage of register prefetch.
// PA / PB point at the expanded nibble arrays.
template <int EPI>
__global__ __launch_bounds__(NTHR, 4) void k_mfma_fp4(MfmaArgs a)
{
    constexpr int KS = 2;               // chunks per LDS stage
    constexpr int LR = KS * 32 + 16;    // LDS bytes per row
    __shared__ __attribute__((aligned(16))) unsigned char As[2][TILE * LR];
    __shared__ __attribute__((aligned(16))) unsigned char Bs[2][TILE * LR];

    const u32 per = (a.ntiles + 7) / 8;
    const u32 t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (t >= a.ntiles) return;
    u32 I, J;
    if (!resolve_tile<EPI>(a, t, I, J)) return;
    const u32 tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const u32 wm = wave >> 1, wn = wave & 1;

    v16f acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[x][y][v] = 0.0f;

    const u32 nst = a.stages;  // 128-site stages
    const bool opB = tid >= TILE;  // wave-uniform: waves 0-1 stage A rows, waves 2-3 B rows
    const u32 row = tid & (TILE - 1);
    const uint4 *E = reinterpret_cast<const uint4 *>(opB ? a.PB : a.PA);
    const size_t pad = opB ? a.padB : a.padA;
    const size_t rbase = (size_t)(opB ? J : I) * TILE + row;
    unsigned char(*Ls)[TILE * LR] = opB ? Bs : As;
    static_assert(KS == 2, "one stage = two 32-byte chunks per row");
    struct Stage {
        uint4 c0a, c0b, c1a, c1b;
    };
    auto ld = [&](u32 s) {
        const uint4 *p0 = E + ((size_t)(s * KS) * pad + rbase) * 2;
        const uint4 *p1 = E + ((size_t)(s * KS + 1) * pad + rbase) * 2;
        return Stage{p0[0], p0[1], p1[0], p1[1]};
    };
    auto st = [&](unsigned char *dst, const Stage &g) {
        uint4 *d = reinterpret_cast<uint4 *>(dst + row * LR);
        d[0] = g.c0a;
        d[1] = g.c0b;
        d[2] = g.c1a;
        d[3] = g.c1b;
    };
    const u32 arow = wm * 64 + (lane & 31), brow = wn * 64 + (lane & 31), koff = 16 * (lane >> 5);
    auto compute = [&](const unsigned char *A, const unsigned char *B) {
#pragma unroll
        for (int c = 0; c < KS; ++c) {
            v8i fa[2], fb[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const v4i ta = *reinterpret_cast<const v4i *>(&A[(arow + 32 * x) * LR + c * 32 + koff]);
                const v4i tb = *reinterpret_cast<const v4i *>(&B[(brow + 32 * x) * LR + c * 32 + koff]);
                fa[x] = v8i{ta.x, ta.y, ta.z, ta.w, 0, 0, 0, 0};
                fb[x] = v8i{tb.x, tb.y, tb.z, tb.w, 0, 0, 0, 0};
            }
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fa[x], fb[y], acc[x][y], 4, 4, 0, 127,
                                                                                0, 127);
        }
    };
    // stage k is held in g0 for even k, g1 for odd k; the loop runs two stages
    // per trip (stages come in multiples of 8) so no register copies are needed,
    // and loads are unconditional (past the end they re-read the last stage)
    // so the compiler can count the loads in flight instead of draining them
    const u32 last = nst - 1;
    Stage g0 = ld(0), g1 = ld(1);
    st(Ls[0], g0);
    g0 = ld(min(2u, last));
    __syncthreads();
    for (u32 s = 0; s < nst; s += 2) {
        compute(As[0], Bs[0]);
        st(Ls[1], g1);
        g1 = ld(min(s + 3, last));
        __syncthreads();
        compute(As[1], Bs[1]);
        st(Ls[0], g0);  // past the end: a stage no one reads
        g0 = ld(min(s + 4, last));
        __syncthreads();
    }

    if constexpr (EPI == E_COUNTS) {
        count_epilogue(a, acc, I, J, wm, wn, lane, tid, reinterpret_cast<u32 *>(&As[0][0]), &Bs[0][0]);
        return;
    }
    const int32_t thr = EPI == E_MASK && a.thr_dev ? *a.thr_dev : a.thr;
    const bool mirror = a.triangle && I != J;
    const u32 er0 = I * TILE - a.orgA + wm * 64, ec0 = J * TILE - a.orgB + wn * 64;  // the group's own rows
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            fp4_epilogue<EPI>(a, acc[x][y], er0 + 32 * x, ec0 + 32 * y, lane, mirror, thr);
        }
}

// fp4 tiles with the B operands read straight from the expanded array into
// registers (each lane's 16-byte operand is contiguous there), so LDS carries
// only the A rows: half the LDS traffic of k_mfma_fp4 for 1.5x its L2 reads.
template <int EPI>
__global__ __launch_bounds__(NTHR, 3) void k_mfma_fp4b(MfmaArgs a)
{
    constexpr int LR = 2 * 32 + 16;  // LDS bytes per A row (one 128-site stage)
    __shared__ __attribute__((aligned(16))) unsigned char As[2][TILE * LR];

    const u32 per = (a.ntiles + 7) / 8;
    const u32 t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (t >= a.ntiles) return;
    u32 I, J;
    if (!resolve_tile<EPI>(a, t, I, J)) return;
    const u32 tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const u32 wm = wave >> 1, wn = wave & 1;

    v16f acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[x][y][v] = 0.0f;

    const u32 nst = a.stages;
    // A staging: thread tid copies chunk (tid >> 7) of row (tid & 127) of each stage
    const u32 arow_s = tid & (TILE - 1), ach = tid >> 7;
    const uint4 *EA = reinterpret_cast<const uint4 *>(a.PA) + ((size_t)I * TILE + arow_s) * 2;
    const size_t padA = a.padA, padB = a.padB;
    auto lda = [&](u32 s, uint4 &p, uint4 &q) {
        const uint4 *src = EA + (size_t)(2 * s + ach) * padA * 2;
        p = src[0];
        q = src[1];
    };
    // B operands: lane's column row of block y, 16 bytes at half (lane >> 5) of each chunk
    const u32 koffB = lane >> 5;
    const uint4 *EB0 = reinterpret_cast<const uint4 *>(a.PB) + ((size_t)J * TILE + wn * 64 + (lane & 31)) * 2 + koffB;
    const uint4 *EB1 = EB0 + 32 * 2;
    auto ldb = [&](u32 s, uint4 &b00, uint4 &b01, uint4 &b10, uint4 &b11) {  // b<chunk><y>
        const size_t o0 = (size_t)(2 * s) * padB * 2, o1 = (size_t)(2 * s + 1) * padB * 2;
        b00 = EB0[o0];
        b01 = EB1[o0];
        b10 = EB0[o1];
        b11 = EB1[o1];
    };
    const u32 arow = wm * 64 + (lane & 31), koff = 16 * (lane >> 5);
    auto stA = [&](unsigned char *dst, const uint4 &p, const uint4 &q) {
        uint4 *d = reinterpret_cast<uint4 *>(&dst[arow_s * LR + ach * 32]);
        d[0] = p;
        d[1] = q;
    };
    auto compute = [&](const unsigned char *A, const uint4 &c00, const uint4 &c01, const uint4 &c10,
                       const uint4 &c11) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            v8i fa[2], fb[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const v4i ta = *reinterpret_cast<const v4i *>(&A[(arow + 32 * x) * LR + c * 32 + koff]);
                fa[x] = v8i{ta.x, ta.y, ta.z, ta.w, 0, 0, 0, 0};
            }
            const uint4 y0 = c ? c10 : c00, y1 = c ? c11 : c01;
            fb[0] = v8i{(int)y0.x, (int)y0.y, (int)y0.z, (int)y0.w, 0, 0, 0, 0};
            fb[1] = v8i{(int)y1.x, (int)y1.y, (int)y1.z, (int)y1.w, 0, 0, 0, 0};
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fa[x], fb[y], acc[x][y], 4, 4, 0, 127,
                                                                                0, 127);
        }
    };
    // two stages per trip (stages come in multiples of 8): B operands of even
    // stages in x**, odd stages in y**; A rows of the next stage in ap / aq
    // loads are unconditional (past the end they re-read the last stage) so the
    // compiler can count the loads in flight instead of draining them
    const u32 last = nst - 1;
    uint4 ap, aq, x00, x01, x10, x11, y00, y01, y10, y11;
    lda(0, ap, aq);
    stA(As[0], ap, aq);
    ldb(0, x00, x01, x10, x11);
    lda(1, ap, aq);
    __syncthreads();
    for (u32 s = 0; s < nst; s += 2) {
        ldb(s + 1, y00, y01, y10, y11);
        compute(As[0], x00, x01, x10, x11);
        stA(As[1], ap, aq);
        lda(min(s + 2, last), ap, aq);
        __syncthreads();
        ldb(min(s + 2, last), x00, x01, x10, x11);
        compute(As[1], y00, y01, y10, y11);
        stA(As[0], ap, aq);  // past the end: a stage no one reads
        lda(min(s + 3, last), ap, aq);
        __syncthreads();
    }

    const int32_t thr = EPI == E_MASK && a.thr_dev ? *a.thr_dev : a.thr;
    const bool mirror = a.triangle && I != J;
    const u32 er0 = I * TILE - a.orgA + wm * 64, ec0 = J * TILE - a.orgB + wn * 64;  // the group's own rows
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            fp4_epilogue<EPI>(a, acc[x][y], er0 + 32 * x, ec0 + 32 * y, lane, mirror, thr);
        }
}

// fp4 tiles fed by LDS DMA: global_load_lds_dwordx4 moves each 16-byte
// granule of a stage (A and B rows, 64 bytes per row per 128-site stage)
// from the expanded array straight into LDS, with no VGPR staging and no
// ds_write.  One DMA instruction fills 1 KB of LDS, lane i at byte 16 i, so
// the lane -> granule map is chosen to give the swizzled row layout the
// operand reads want: row r holds granules 0..3 (chunk c, half h at g = 2c + h)
// at positions g ^ ((r >> 2) & 3), which keeps 16 consecutive rows of one
// granule on distinct banks (a 16-byte read serves 16 lanes per cycle from
// 64 banks: rows r, r + 4, r + 8, r + 12 share a bank window of 64-byte rows
// and take its 4 positions).  Two stage buffers (separate LDS objects, loop
// unrolled by two): the DMA of stage s + 1 runs under the MFMAs of stage s.
constexpr u32 FP4D_STAGE = 2 * TILE * 64;  // bytes of one stage buffer (A then B)

__device__ __forceinline__ void fp4d_dma(unsigned char *buf, const unsigned char *EA, const unsigned char *EB,
                                         size_t padA, size_t padB, u32 I, u32 J, u32 s, u32 wave, u32 lane)
{
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const u32 q = (wave * 4 + k) * 64 + lane;  // granule index in the stage buffer
        const bool opB = q >= TILE * 4;
        const u32 q2 = q & (TILE * 4 - 1), row = q2 >> 2, g = (q2 & 3) ^ ((row >> 2) & 3);
        const unsigned char *E = opB ? EB : EA;
        const size_t pad = opB ? padB : padA;
        const size_t rr = (size_t)(opB ? J : I) * TILE + row;
        const unsigned char *src = E + ((size_t)(2 * s + (g >> 1)) * pad + rr) * 32 + (g & 1) * 16;
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                         (__attribute__((address_space(3))) void *)(buf + (wave * 4 + k) * 1024),
                                         16, 0, 0);
    }
}

template <int EPI>
__global__ __launch_bounds__(NTHR, 4) void k_mfma_fp4d(MfmaArgs a)
{
    __shared__ __attribute__((aligned(16))) unsigned char S0[FP4D_STAGE];
    __shared__ __attribute__((aligned(16))) unsigned char S1[FP4D_STAGE];

    const u32 per = (a.ntiles + 7) / 8;
    const u32 t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (t >= a.ntiles) return;
    u32 I, J;
    if (!resolve_tile<EPI>(a, t, I, J)) return;
    const u32 tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const u32 wm = wave >> 1, wn = wave & 1;

    v16f acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[x][y][v] = 0.0f;

    const u32 nst = a.stages, last = nst - 1;
    const unsigned char *EA = reinterpret_cast<const unsigned char *>(a.PA);
    const unsigned char *EB = reinterpret_cast<const unsigned char *>(a.PB);
    const size_t padA = a.padA, padB = a.padB;
    const u32 ra = wm * 64 + (lane & 31), rb = wn * 64 + (lane & 31), h = lane >> 5;
    auto compute = [&](const unsigned char *buf) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            v8i fa[2], fb[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const u32 r1 = ra + 32 * x, r2 = rb + 32 * x;
                const u32 g = 2 * c + h;
                const v4i ta = *reinterpret_cast<const v4i *>(buf + r1 * 64 + ((g ^ ((r1 >> 2) & 3)) * 16));
                const v4i tb =
                    *reinterpret_cast<const v4i *>(buf + TILE * 64 + r2 * 64 + ((g ^ ((r2 >> 2) & 3)) * 16));
                fa[x] = v8i{ta.x, ta.y, ta.z, ta.w, 0, 0, 0, 0};
                fb[x] = v8i{tb.x, tb.y, tb.z, tb.w, 0, 0, 0, 0};
            }
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fa[x], fb[y], acc[x][y], 4, 4, 0, 127,
                                                                                0, 127);
        }
    };
    // stages come in multiples of 8; past the end the DMA re-reads the last
    // stage into a buffer no one reads again
    // sched_barrier keeps each DMA issue ahead of the MFMAs it overlaps
    fp4d_dma(S0, EA, EB, padA, padB, I, J, 0, wave, lane);
    for (u32 s = 0; s < nst; s += 2) {
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's DMA of stage s has landed
        __syncthreads();                     // ... and every wave's
        fp4d_dma(S1, EA, EB, padA, padB, I, J, s + 1, wave, lane);
        __builtin_amdgcn_sched_barrier(0);
        compute(S0);
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
        fp4d_dma(S0, EA, EB, padA, padB, I, J, min(s + 2, last), wave, lane);
        __builtin_amdgcn_sched_barrier(0);
        compute(S1);
        __builtin_amdgcn_sched_barrier(0);
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);

    const int32_t thr = EPI == E_MASK && a.thr_dev ? *a.thr_dev : a.thr;
    const bool mirror = a.triangle && I != J;
    const u32 er0 = I * TILE - a.orgA + wm * 64, ec0 = J * TILE - a.orgB + wn * 64;  // the group's own rows
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            fp4_epilogue<EPI>(a, acc[x][y], er0 + 32 * x, ec0 + 32 * y, lane, mirror, thr);
        }
}

// fp4 with 256 x 256 pair tiles: 8 waves (2 x 4), each 128 x 64 pairs as
// 4 x 2 MFMA blocks, stages moved by LDS DMA as in k_mfma_fp4d (same swizzled
// 64-byte rows, A rows then B rows).  Per pair it reads half the bytes of the
// 128 x 128 tiles, which is what bounds them (each CU streams its tiles' rows
// from L2 at the rate its MFMAs consume them).
constexpr u32 FP4W_STAGE = 2 * WT * 64;  // bytes of one stage buffer

__device__ __forceinline__ void fp4w_dma(unsigned char *buf, const unsigned char *EA, const unsigned char *EB,
                                         size_t padA, size_t padB, u32 I, u32 J, u32 s, u32 wave, u32 lane)
{
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const u32 q = (wave * 4 + k) * 64 + lane;  // granule index in the stage buffer
        const bool opB = q >= WT * 4;
        const u32 q2 = q & (WT * 4 - 1), row = q2 >> 2, g = (q2 & 3) ^ ((row >> 2) & 3);
        const unsigned char *E = opB ? EB : EA;
        const size_t pad = opB ? padB : padA;
        const size_t rr = (size_t)(opB ? J : I) * WT + row;
        const unsigned char *src = E + ((size_t)(2 * s + (g >> 1)) * pad + rr) * 32 + (g & 1) * 16;
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                         (__attribute__((address_space(3))) void *)(buf + (wave * 4 + k) * 1024),
                                         16, 0, 0);
    }
}

template <int EPI>
__global__ __launch_bounds__(512, 1) void k_mfma_fp4w(MfmaArgs a)
{
    __shared__ __attribute__((aligned(16))) unsigned char S0[FP4W_STAGE];
    __shared__ __attribute__((aligned(16))) unsigned char S1[FP4W_STAGE];

    const u32 per = (a.ntiles + 7) / 8;
    const u32 t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (t >= a.ntiles || mask_skipped<EPI>(a)) return;
    u32 I, J;
    if (!tile_of(a, t, I, J)) return;
    const u32 tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const u32 wm = wave >> 2, wn = wave & 3;

    v16f acc[4][2];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[x][y][v] = 0.0f;

    const u32 nst = a.stages, last = nst - 1;
    const unsigned char *EA = reinterpret_cast<const unsigned char *>(a.PA);
    const unsigned char *EB = reinterpret_cast<const unsigned char *>(a.PB);
    const size_t padA = a.padA, padB = a.padB;
    const u32 ra = wm * 128 + (lane & 31), rb = wn * 64 + (lane & 31), h = lane >> 5;
    auto compute = [&](const unsigned char *buf) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const u32 g = 2 * c + h;
            v8i fa[4], fb[2];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                const u32 r1 = ra + 32 * x;
                const v4i ta = *reinterpret_cast<const v4i *>(buf + r1 * 64 + ((g ^ ((r1 >> 2) & 3)) * 16));
                fa[x] = v8i{ta.x, ta.y, ta.z, ta.w, 0, 0, 0, 0};
            }
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const u32 r2 = rb + 32 * y;
                const v4i tb = *reinterpret_cast<const v4i *>(buf + WT * 64 + r2 * 64 + ((g ^ ((r2 >> 2) & 3)) * 16));
                fb[y] = v8i{tb.x, tb.y, tb.z, tb.w, 0, 0, 0, 0};
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fa[x], fb[y], acc[x][y], 4, 4, 0, 127,
                                                                                0, 127);
        }
    };
    fp4w_dma(S0, EA, EB, padA, padB, I, J, 0, wave, lane);
    for (u32 s = 0; s < nst; s += 2) {
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's DMA of stage s has landed
        __syncthreads();                     // ... and every wave's
        fp4w_dma(S1, EA, EB, padA, padB, I, J, s + 1, wave, lane);
        __builtin_amdgcn_sched_barrier(0);
        compute(S0);
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
        fp4w_dma(S0, EA, EB, padA, padB, I, J, min(s + 2, last), wave, lane);
        __builtin_amdgcn_sched_barrier(0);
        compute(S1);
        __builtin_amdgcn_sched_barrier(0);
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);

    const int32_t thr = EPI == E_MASK && a.thr_dev ? *a.thr_dev : a.thr;
    const bool mirror = a.triangle && I != J;
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            fp4_epilogue<EPI>(a, acc[x][y], I * WT + wm * 128 + 32 * x, J * WT + wn * 64 + 32 * y, lane, mirror, thr);
        }
}

// fp4 with 256 x 256 pair tiles of 4 waves, one per SIMD, each 128 x 128 pairs
// as 4 x 4 MFMA blocks (256 fp32 accumulators per lane): per pair a quarter of
// the LDS fragment reads of the 128 x 128 tiles and half their L2 reads.  With
// no second wave on its SIMD, a wave keeps its matrix core fed by software
// pipelining instead of occupancy:
//   * four LDS stage buffers (the swizzled 64-byte rows of k_mfma_fp4w), filled
//     by LDS DMA three stages ahead of the stage being multiplied;
//   * the fragments of a stage's second 64-site chunk are read while the first
//     chunk's 16 MFMAs run (two register sets of 8 fragments);
//   * the wait for the next stage's DMA and the barrier that publishes it come
//     before the last 4 MFMAs of a stage, so the next stage's first fragment
//     reads overlap them.
// The buffers are four arrays used in a fixed order (the loop is unrolled by
// four), so the compiler sees which buffer each read and DMA touches.
constexpr u32 FP4X_STAGE = 2 * WT * 64;  // bytes of one stage buffer: A rows then B rows

template <int EPI>
__global__ __launch_bounds__(256, 1) void k_mfma_fp4x(MfmaArgs a)
{
    __shared__ __attribute__((aligned(16))) unsigned char S0[FP4X_STAGE];
    __shared__ __attribute__((aligned(16))) unsigned char S1[FP4X_STAGE];
    __shared__ __attribute__((aligned(16))) unsigned char S2[FP4X_STAGE];
    __shared__ __attribute__((aligned(16))) unsigned char S3[FP4X_STAGE];

    const u32 per = (a.ntiles + 7) / 8;
    const u32 t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (t >= a.ntiles || mask_skipped<EPI>(a)) return;
    u32 I, J;
    if (!tile_of(a, t, I, J)) return;
    const u32 tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const u32 wm = wave >> 1, wn = wave & 1, h = lane >> 5;

    v16f acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[x][y][v] = 0.0f;

    // DMA: waves 0-1 move the A rows, waves 2-3 the B rows; lane granule q of
    // instruction k is slot q & 3 of row q >> 2 and holds granule
    // g = slot ^ ((row >> 2) & 3) (chunk g >> 1, half g & 1) of that row
    const bool opB = wave >= 2;
    const unsigned char *E = reinterpret_cast<const unsigned char *>(opB ? a.PB : a.PA);
    const size_t pad = opB ? a.padB : a.padA;
    const size_t rbase = (size_t)(opB ? J : I) * WT;
    size_t off[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const u32 q = ((wave & 1) * 8 + k) * 64 + lane, row = q >> 2, g = (q & 3) ^ ((row >> 2) & 3);
        off[k] = ((size_t)(g >> 1) * pad + rbase + row) * 32 + (g & 1) * 16;
    }
    const size_t sstride = 64 * pad;  // bytes between stages: two 32-byte chunks per row
    const u32 dbase = wave * 8 * 1024;
    auto dma = [&](unsigned char *buf, u32 s) {
        const unsigned char *src = E + (size_t)s * sstride;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(src + off[k]),
                                             (__attribute__((address_space(3))) void *)(buf + dbase + k * 1024), 16,
                                             0, 0);
    };
    // fragment of row r, chunk c: the row's slot (2c + h) ^ ((r >> 2) & 3); with
    // r = 32 m + (lane & 31) the swizzle depends on the lane only
    const u32 lo0 = (lane & 31) * 64 + ((h ^ ((lane >> 2) & 3)) * 16), lo1 = lo0 ^ 32;
    const u32 fa0 = wm * 128 * 64, fb0 = WT * 64 + wn * 128 * 64;
    v4i F0[8], F1[8];
    auto frags = [&](const unsigned char *buf, u32 lo, v4i(&F)[8]) {
#pragma unroll
        for (int x = 0; x < 4; ++x) F[x] = *reinterpret_cast<const v4i *>(buf + fa0 + x * 32 * 64 + lo);
#pragma unroll
        for (int y = 0; y < 4; ++y) F[4 + y] = *reinterpret_cast<const v4i *>(buf + fb0 + y * 32 * 64 + lo);
    };
    auto mma = [&](const v4i(&F)[8], int b) {
        const int x = b >> 2, y = b & 3;
        const v8i fa = v8i{F[x].x, F[x].y, F[x].z, F[x].w, 0, 0, 0, 0};
        const v8i fb = v8i{F[4 + y].x, F[4 + y].y, F[4 + y].z, F[4 + y].w, 0, 0, 0, 0};
        acc[x][y] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fa, fb, acc[x][y], 4, 4, 0, 127, 0, 127);
    };
    // one stage out of buffer `cur`: PRE issues the DMA of stage s + 3 into
    // `pre`, NEXT publishes stage s + 1 (buffer `nxt`) and reads its first
    // fragments.  Waits: before the barrier this wave's DMA of stage s + 1 has
    // landed -- the DMAs of s + 2 and s + 3 (16 instructions) may still be in
    // flight when PRE, none otherwise.
    auto step = [&](u32 s, unsigned char *cur, unsigned char *nxt, unsigned char *pre, auto PRE, auto NEXT) {
        frags(cur, lo1, F1);
        if constexpr (decltype(PRE)::value) dma(pre, s + 3);
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            mma(F0, b);
            if (b < 8) __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
            if (b < 8) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // LDS read
            if (b >= 8) __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            if (b >= 8 && decltype(PRE)::value) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // DMA
        }
        __builtin_amdgcn_sched_barrier(0);  // the second chunk's MFMAs stay behind the first's
#pragma unroll
        for (int b = 0; b < 12; ++b) mma(F1, b);
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (decltype(NEXT)::value) {
            if constexpr (decltype(PRE)::value)
                __builtin_amdgcn_s_waitcnt(0x4F70);  // vmcnt(16)
            else
                __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
            asm volatile("" ::: "memory");  // no compiler motion of LDS accesses across the barrier
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
            frags(nxt, lo0, F0);
        }
#pragma unroll
        for (int b = 12; b < 16; ++b) {
            mma(F1, b);
            if constexpr (decltype(NEXT)::value) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 1);
                __builtin_amdgcn_sched_group_barrier(0x100, 2, 1);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    using T_ = std::true_type;
    using F_ = std::false_type;
    // stages come in multiples of 8
    const u32 nst = a.stages;
    dma(S0, 0);
    dma(S1, 1);
    dma(S2, 2);
    __builtin_amdgcn_s_waitcnt(0x4F70);  // vmcnt(16): stage 0 has landed
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    frags(S0, lo0, F0);
    u32 s = 0;
    for (; s + 4 < nst; s += 4) {
        step(s, S0, S1, S3, T_{}, T_{});
        step(s + 1, S1, S2, S0, T_{}, T_{});
        step(s + 2, S2, S3, S1, T_{}, T_{});
        step(s + 3, S3, S0, S2, T_{}, T_{});
    }
    step(s, S0, S1, S3, T_{}, T_{});
    step(s + 1, S1, S2, S0, F_{}, T_{});
    step(s + 2, S2, S3, S1, F_{}, T_{});
    step(s + 3, S3, S0, S2, F_{}, F_{});

    const int32_t thr = EPI == E_MASK && a.thr_dev ? *a.thr_dev : a.thr;
    const bool mirror = a.triangle && I != J;
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            fp4_epilogue<EPI>(a, acc[x][y], I * WT + wm * 128 + 32 * x, J * WT + wn * 128 + 32 * y, lane, mirror, thr);
        }
}

// fp4 with 256 x 128 pair tiles (256 A rows, 128 B rows) of 4 waves, each
// 128 x 64 pairs as 4 x 2 MFMA blocks (128 fp32 accumulators per lane), so two
// workgroups share a CU: while one runs its epilogue (about 100 VALU
// instructions per 32 x 32 block for the radius mask) or waits for its first
// stage, the other's MFMAs keep the matrix cores busy.  Inside a workgroup the
// pipeline of k_mfma_fp4x: LDS DMA two stages ahead into three rotating
// buffers (24 KB each: A rows then B rows, swizzled 64-byte rows), the second
// chunk's fragments read during the first chunk's MFMAs, the next stage
// published before the last two MFMAs of a stage.  Triangles run on 256-row blocks of A
// against the two 128-row halves of the matching 256-row blocks of B; a 32 x 32
// block below the diagonal is skipped (its mirror is written by the block above
// it) and one on the diagonal writes its own pairs only.
constexpr u32 FP4Y_STAGE = (WT + TILE) * 64;  // 24 KB

template <int EPI>
__global__ __launch_bounds__(256, 2) void k_mfma_fp4y(MfmaArgs a)
{
    __shared__ __attribute__((aligned(16))) unsigned char S[3 * FP4Y_STAGE];

    const u32 per = (a.ntiles + 7) / 8;
    const u32 t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (t >= a.ntiles || mask_skipped<EPI>(a)) return;
    u32 I, J;  // A block of 256 rows, B block of 128 rows
    if (a.triangle) {
        u32 I2, J2;
        tri_index(t >> 1, I2, J2);
        I = I2;
        J = 2 * J2 + (t & 1);
        if (J >= a.ntB) return;  // the second half of the last 256-row block of B
    } else {
        I = t / a.ntB;
        J = t - I * a.ntB;
    }
    const u32 tid = threadIdx.x, lane = tid & 63;
    const u32 wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: the DMA bases stay scalar
    const u32 wm = wave >> 1, wn = wave & 1, h = lane >> 5;

    v16f acc[4][2];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[x][y][v] = 0.0f;

    // DMA: 1,536 granules per stage, 6 instructions per wave; instruction
    // w * 6 + k covers granules 64 (w * 6 + k) .. + 63, A rows (granules below
    // 1,024) or B rows, uniformly per instruction; granule q of a row block is
    // slot q & 3 of row q >> 2 and holds granule g = slot ^ ((row >> 2) & 3)
    // the A rows of the tile start at EA + 32 I * 256, the B rows at EB + 32 J * 128
    // (chunk 0); a lane's offset from there is below 64 pad bytes
    const unsigned char *baseA = reinterpret_cast<const unsigned char *>(a.PA) + (size_t)I * WT * 32;
    const unsigned char *baseB = reinterpret_cast<const unsigned char *>(a.PB) + (size_t)J * TILE * 32;
    const size_t strA = 64 * (size_t)a.padA, strB = 64 * (size_t)a.padB;  // bytes per stage
    u32 off[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const bool opB = wave * 6 + k >= 16;  // uniform
        const u32 q2 = (wave * 6 + k) * 64 + lane - (opB ? WT * 4 : 0u), row = q2 >> 2;
        const u32 g = (q2 & 3) ^ ((row >> 2) & 3);
        off[k] = ((g >> 1) * (opB ? a.padB : a.padA) + row) * 32 + (g & 1) * 16;
    }
    const u32 dbase = wave * 6 * 1024;
    auto dma = [&](unsigned char *buf, u32 s) {
        const unsigned char *sa = baseA + s * strA, *sb = baseB + s * strB;
#pragma unroll
        for (int k = 0; k < 6; ++k)
            __builtin_amdgcn_global_load_lds(
                (const __attribute__((address_space(1))) void *)((wave * 6 + k >= 16 ? sb : sa) + off[k]),
                (__attribute__((address_space(3))) void *)(buf + dbase + k * 1024), 16, 0, 0);
    };
    const u32 lo0 = (lane & 31) * 64 + ((h ^ ((lane >> 2) & 3)) * 16), lo1 = lo0 ^ 32;
    const u32 fa0 = wm * 128 * 64, fb0 = WT * 64 + wn * 64 * 64;
    v4i F0[6], F1[6];
    auto frags = [&](const unsigned char *buf, u32 lo, v4i(&F)[6]) {
#pragma unroll
        for (int x = 0; x < 4; ++x) F[x] = *reinterpret_cast<const v4i *>(buf + fa0 + x * 32 * 64 + lo);
#pragma unroll
        for (int y = 0; y < 2; ++y) F[4 + y] = *reinterpret_cast<const v4i *>(buf + fb0 + y * 32 * 64 + lo);
    };
    auto mma = [&](const v4i(&F)[6], int b) {
        const int x = b >> 1, y = b & 1;
        const v8i fa = v8i{F[x].x, F[x].y, F[x].z, F[x].w, 0, 0, 0, 0};
        const v8i fb = v8i{F[4 + y].x, F[4 + y].y, F[4 + y].z, F[4 + y].w, 0, 0, 0, 0};
        acc[x][y] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fa, fb, acc[x][y], 4, 4, 0, 127, 0, 127);
    };
    // one stage out of buffer `cur`: the DMA of stage s + 2 (the last stage
    // again past the end, into a buffer no one reads) goes into `pre`; then
    // stage s + 1 (in `nxt`) is published -- this wave's DMA of it has landed
    // (vmcnt(6) leaves the DMA of s + 2 in flight) and every wave has passed
    // the barrier -- and its first fragments are read.  The three buffers
    // rotate, so every stage runs the same code.
    const u32 nst = a.stages, last = nst - 1;
    dma(S, 0);
    dma(S + FP4Y_STAGE, 1);
    __builtin_amdgcn_s_waitcnt(0x0F76);  // vmcnt(6): stage 0 has landed
    asm volatile("" ::: "memory");  // no compiler motion of LDS accesses across the barrier
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    frags(S, lo0, F0);
    u32 cur = 0, nxt = FP4Y_STAGE, pre = 2 * FP4Y_STAGE;
    for (u32 s = 0; s < nst; ++s) {
        frags(S + cur, lo1, F1);
        dma(S + pre, min(s + 2, last));
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            mma(F0, b);
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
            if (b < 6) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // LDS read
            if (b >= 2) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // DMA
        }
        __builtin_amdgcn_sched_barrier(0);  // the second chunk's MFMAs stay behind the first's
#pragma unroll
        for (int b = 0; b < 6; ++b) mma(F1, b);
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(0x0F76);  // vmcnt(6)
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
        frags(S + nxt, lo0, F0);
#pragma unroll
        for (int b = 6; b < 8; ++b) {
            mma(F1, b);
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 1);
            __builtin_amdgcn_sched_group_barrier(0x100, 3, 1);
        }
        __builtin_amdgcn_sched_barrier(0);
        const u32 c = cur;
        cur = nxt;
        nxt = pre;
        pre = c;
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): no DMA still writes this workgroup's LDS

    const int32_t thr = EPI == E_MASK && a.thr_dev ? *a.thr_dev : a.thr;
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const u32 r0 = I * WT + wm * 128 + 32 * x, c0 = J * TILE + wn * 64 + 32 * y;
            if (a.triangle && c0 < r0) continue;  // wave-uniform
            fp4_epilogue<EPI>(a, acc[x][y], r0, c0, lane, a.triangle && c0 > r0, thr);
        }
}

// fp4 with 256 x 256 pair tiles and one persistent workgroup per CU.  Each
// XCD (blockIdx.x & 7) owns a contiguous run of tiles, which its workgroups
// take round-robin; the LDS DMA never drains between tiles: a workgroup's next
// tile streams its first four stages in during the last four stages of the
// current tile and its epilogue.  A 256 x 256 tile moves 32 KB per stage
// (128 sites) for 65,536 pairs, against 48 KB for the two 256 x 128 tiles of
// k_mfma_fp4y, and the DMA into LDS is what bounds those tiles.  Four waves of
// 128 x 128 pairs (4 x 4 MFMA blocks, 256 fp32 accumulators per lane); five
// rotating LDS buffers of 32 KB (A rows then B rows, swizzled 64-byte rows:
// the whole 160 KB), DMA four stages ahead.  Per stage: the second chunk's fragments and the
// stage's DMA are issued under the first chunk's 16 MFMAs, then the barrier
// (this wave's DMA of the next stage has landed and its reads of this stage
// are done), then the next stage's first fragments are read under the second
// chunk's 16 MFMAs.
constexpr u32 FP4P_STAGE = 2 * WT * 64;  // 32 KB
constexpr u32 FP4P_NBUF = 5;

template <int EPI>
__global__ __launch_bounds__(256, 1) void k_mfma_fp4p(MfmaArgs a)
{
    __shared__ __attribute__((aligned(16))) unsigned char S[FP4P_NBUF * FP4P_STAGE];

    if (mask_skipped<EPI>(a)) return;
    const u32 per = (a.ntiles + 7) / 8, xcd = blockIdx.x & 7;
    const u32 tend = min((xcd + 1) * per, a.ntiles), tstep = gridDim.x >> 3;
    // (slot, I, J) of the first tile at or after slot t, slot tend when none
    auto next_tile = [&](u32 t) {
        for (; t < tend; t += tstep) {
            u32 i, j;
            if (tile_of(a, t, i, j)) return make_uint3(t, i, j);
        }
        return make_uint3(tend, 0u, 0u);
    };
    uint3 tile = next_tile(xcd * per + (blockIdx.x >> 3));
    if (tile.x >= tend) return;

    const u32 tid = threadIdx.x, lane = tid & 63;
    const u32 wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: the DMA bases stay scalar
    const u32 wm = wave >> 1, wn = wave & 1, h = lane >> 5;

    // DMA: waves 0-1 move the A rows, waves 2-3 the B rows, 8 instructions of
    // 64 x 16 bytes each; lane granule q of instruction k is slot q & 3 of row
    // q >> 2 and holds granule g = slot ^ ((row >> 2) & 3) (chunk g >> 1, half
    // g & 1) of that row
    const bool opB = wave >= 2;
    const unsigned char *E = reinterpret_cast<const unsigned char *>(opB ? a.PB : a.PA);
    const u32 pad = opB ? a.padB : a.padA;  // below 2^26 (checked on the host)
    u32 off[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const u32 q = ((wave & 1) * 8 + k) * 64 + lane, row = q >> 2, g = (q & 3) ^ ((row >> 2) & 3);
        off[k] = ((g >> 1) * pad + row) * 32 + (g & 1) * 16;
    }
    const size_t sstride = 64 * (size_t)pad;  // bytes between stages: two 32-byte chunks per row
    const u32 dbase = wave * 8 * 1024;
    auto base_of = [&](const uint3 &tl) {
        return E + (size_t)__builtin_amdgcn_readfirstlane(opB ? tl.z : tl.y) * WT * 32;
    };
    auto dma = [&](u32 buf, const unsigned char *base, u32 st) {
        const unsigned char *src = base + (size_t)st * sstride;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(src + off[k]),
                                             (__attribute__((address_space(3))) void *)(S + buf + dbase + k * 1024),
                                             16, 0, 0);
    };
    // fragment of row r, chunk c: the row's slot (2c + h) ^ ((r >> 2) & 3); with
    // r = 32 m + (lane & 31) the swizzle depends on the lane only
    const u32 lo0 = (lane & 31) * 64 + ((h ^ ((lane >> 2) & 3)) * 16), lo1 = lo0 ^ 32;
    const u32 fa0 = wm * 128 * 64, fb0 = WT * 64 + wn * 128 * 64;
    v4i F0[8], F1[8];
    auto frags = [&](u32 buf, u32 lo, v4i(&F)[8]) {
#pragma unroll
        for (int x = 0; x < 4; ++x) F[x] = *reinterpret_cast<const v4i *>(S + buf + fa0 + x * 32 * 64 + lo);
#pragma unroll
        for (int y = 0; y < 4; ++y) F[4 + y] = *reinterpret_cast<const v4i *>(S + buf + fb0 + y * 32 * 64 + lo);
    };
    v16f acc[4][4];
    // FIRST: the tile's first MFMA of every block, which takes its
    // accumulator input from block (3, 3), zeroed before and itself last
    auto mma = [&](const v4i(&F)[8], int b, auto FIRST) {
        const int x = b >> 2, y = b & 3;
        const v8i fa = v8i{F[x].x, F[x].y, F[x].z, F[x].w, 0, 0, 0, 0};
        const v8i fb = v8i{F[4 + y].x, F[4 + y].y, F[4 + y].z, F[4 + y].w, 0, 0, 0, 0};
        acc[x][y] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fa, fb, decltype(FIRST)::value ? acc[3][3] : acc[x][y],
                                                                  4, 4, 0, 127, 0, 127);
    };
    // one stage out of buffer `cur`; its DMA (base db, stage sd) goes into `pre`;
    // NEXT: the next stage's first fragments are read (the last stage of a
    // tile leaves them until after the epilogue, whose registers they would
    // crowd)
    auto stage = [&](u32 cur, u32 nxt, u32 pre, const unsigned char *db, u32 sd, auto NEXT, auto FIRST) {
        frags(cur, lo1, F1);
        dma(pre, db, sd);
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            mma(F0, b, FIRST);
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
            if (b < 8) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // LDS read
            if (b >= 8) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // DMA
        }
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(0x4078);  // vmcnt(24) lgkmcnt(0): the next stage's DMA has landed
        asm volatile("" ::: "memory");  // no compiler motion of LDS accesses across the barrier
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (decltype(NEXT)::value) frags(nxt, lo0, F0);
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            mma(F1, b, std::false_type{});
            if (decltype(NEXT)::value && b < 8) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 1);
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 1);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    };

    const u32 nst = a.stages;  // at least 4 (checked on the host)
    auto bnext = [](u32 b) { return b + 1 == FP4P_NBUF ? 0u : b + 1; };
    const int32_t thr = EPI == E_MASK && a.thr_dev ? *a.thr_dev : a.thr;
    const unsigned char *base = base_of(tile);
    dma(0, base, 0);
    dma(FP4P_STAGE, base, 1);
    dma(2 * FP4P_STAGE, base, 2);
    dma(3 * FP4P_STAGE, base, 3);
    __builtin_amdgcn_s_waitcnt(0x4F78);  // vmcnt(24): stage 0 has landed
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    frags(0, lo0, F0);
    u32 cur = 0;
    for (;;) {
        const uint3 nt = next_tile(tile.x + tstep);
        const bool more = nt.x < tend;
        const unsigned char *nbase = more ? base_of(nt) : base;
        // stage s prefetches stage s + 4 of this tile or, past its end, the
        // next tile's first stages (the last workgroup tile reloads its own
        // last stage into a buffer no one reads)
        // block (3, 3) is zeroed in its registers, out of the compiler's
        // sight: a known zero becomes the inline-constant accumulator input
        // of the tile's first MFMAs, and with the accumulators in AGPRs that
        // form gave wrong sums on this target (a one-wave check with VGPR
        // accumulators gave the right ones)
        acc[3][3] = (v16f)(0.0f);
        asm volatile("" : "+a"(acc[3][3]));
        auto run = [&](u32 s, auto NEXT, auto FIRST) {
            const u32 sd = s + 4, nx = bnext(cur), pre = cur == 0 ? FP4P_NBUF - 1 : cur - 1;
            const bool here = sd < nst;
            stage(cur * FP4P_STAGE, nx * FP4P_STAGE, pre * FP4P_STAGE, here ? base : nbase,
                  here ? sd : (more ? sd - nst : nst - 1), NEXT, FIRST);
            cur = nx;
        };
        run(0, std::true_type{}, std::true_type{});
        for (u32 s = 1; s + 1 < nst; ++s) run(s, std::true_type{}, std::false_type{});
        run(nst - 1, std::false_type{}, std::false_type{});
        const bool mirror = a.triangle && tile.y != tile.z;
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                fp4_epilogue<EPI>(a, acc[x][y], tile.y * WT + wm * 128 + 32 * x, tile.z * WT + wn * 128 + 32 * y, lane,
                                  mirror, thr);
                __builtin_amdgcn_sched_barrier(0);  // one block's registers at a time
            }
        if (!more) break;
        // the next tile's stage 0 landed before the last stage's barrier and
        // no DMA writes its buffer before that tile's second stage
        frags(cur * FP4P_STAGE, lo0, F0);
        tile = nt;
        base = nbase;
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): no DMA still writes this workgroup's LDS
}

__global__ void k_stage_popc(const u32 *P, u32 n, u32 stride, u32 stages, u32 *pop)
{
    const u32 r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    u32 c = 0;
    for (u32 st = 0; st < stages; ++st) {
        const uint4 v = reinterpret_cast<const uint4 *>(P)[(size_t)st * stride + r];
        c += __builtin_popcount(v.x) + __builtin_popcount(v.y) + __builtin_popcount(v.z) + __builtin_popcount(v.w);
    }
    pop[r] = c;
}

u32 grid_for(u64 items) { return (u32)std::min<u64>((items + 255) / 256, 65535ull * 8); }

}  // namespace

void launch_to_stage_major(const u32 *mat, u32 n, u32 ldw, u32 kw, u32 *P, hipStream_t s, const u32 *rowidx,
                           u32 nsrc)
{
    const u32 nrows_pad = stage_rows(n);
    hipLaunchKernelGGL(k_to_stage_major, dim3(grid_for((u64)(kw / 4) * nrows_pad)), dim3(256), 0, s, mat, rowidx, n,
                       ldw, kw, nrows_pad, P, nsrc);
    CASPHY_HIP(hipGetLastError());
}

void launch_expand_fp4(const u32 *P, u32 pad, u32 stages, u32 *E, hipStream_t s)
{
    hipLaunchKernelGGL(k_expand_fp4, dim3(grid_for((u64)stages * pad)), dim3(256), 0, s, P, pad, stages, E);
    CASPHY_HIP(hipGetLastError());
}

void launch_stage_popc(const u32 *P, u32 n, u32 stride, u32 ldw, u32 *pop, hipStream_t s)
{
    hipLaunchKernelGGL(k_stage_popc, dim3(ceil_div(n, 256)), dim3(256), 0, s, P, n, stride, ldw / 4, pop);
    CASPHY_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Register-resident variant: no LDS.  A workgroup of 4 independent waves owns a
// 256 x 256 pair tile; each wave owns 128 x 128 pairs as 4 x 4 MFMA blocks
// (256 int32 accumulators per lane).  Per 128-site stage a lane loads one
// 16-byte stage-major chunk for each of its 8 row blocks (32 consecutive rows
// per block: one 512-byte coalesced piece; the two lane halves share it) and
// expands only the 16 bits its MFMA operand needs, right before the 16 MFMAs
// that use them.  Each expanded fragment feeds 4 MFMAs.
constexpr u32 RT = 256;  // workgroup tile

template <int EPI>
__global__ __launch_bounds__(256, 1) void k_mfma_reg(MfmaArgs a)
{
    const u32 per = (a.ntiles + 7) / 8;
    const u32 t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (t >= a.ntiles || mask_skipped<EPI>(a)) return;
    u32 I, J;
    if (!tile_of(a, t, I, J)) return;
    const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const u32 wm = wave >> 1, wn = wave & 1;
    const u32 half = lane >> 5, col = lane & 31;
    const int32_t thr = EPI == E_MASK && a.thr_dev ? *a.thr_dev : a.thr;
    const u32 arow0 = I * RT + wm * 128, brow0 = J * RT + wn * 128;
    const uint4 *PA = reinterpret_cast<const uint4 *>(a.PA) + arow0 + col;
    const uint4 *PB = reinterpret_cast<const uint4 *>(a.PB) + brow0 + col;

    v16i acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[x][y][v] = 0;

    uint4 ra[4], rb[4], na[4], nb[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        ra[x] = PA[32 * x];
        rb[x] = PB[32 * x];
    }
    const u32 nst = a.stages;
    for (u32 s = 0; s < nst; ++s) {
        if (s + 1 < nst) {
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                na[x] = PA[(size_t)(s + 1) * a.padA + 32 * x];
                nb[x] = PB[(size_t)(s + 1) * a.padB + 32 * x];
            }
        }
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            v4i fa[4], fb[4];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                const u32 wa = kk == 0 ? ra[x].x : kk == 1 ? ra[x].y : kk == 2 ? ra[x].z : ra[x].w;
                const u32 wb = kk == 0 ? rb[x].x : kk == 1 ? rb[x].y : kk == 2 ? rb[x].z : rb[x].w;
                const uint4 ea = spread16(half ? wa >> 16 : wa);
                const uint4 eb = spread16(half ? wb >> 16 : wb);
                fa[x] = v4i{(int)ea.x, (int)ea.y, (int)ea.z, (int)ea.w};
                fb[x] = v4i{(int)eb.x, (int)eb.y, (int)eb.z, (int)eb.w};
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[x], fb[y], acc[x][y], 0, 0, 0);
        }
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            ra[x] = na[x];
            rb[x] = nb[x];
        }
    }

    const bool mirror = a.triangle && I != J;
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) block_epilogue<EPI>(a, acc[x][y], arow0 + 32 * x, brow0 + 32 * y, lane, mirror, thr);
}

template <int KB, int PF, int OCC>
void launch_variant(int epi, const MfmaArgs &a, hipStream_t s)
{
    const u32 grid = 8 * ((a.ntiles + 7) / 8);
    if (epi == E_MASK)
        hipLaunchKernelGGL((k_mfma_tiles<E_MASK, KB, PF, OCC>), dim3(grid), dim3(NTHR), 0, s, a);
    else if (epi == E_SCORES)
        hipLaunchKernelGGL((k_mfma_tiles<E_SCORES, KB, PF, OCC>), dim3(grid), dim3(NTHR), 0, s, a);
    else if (epi == E_LOCALMAX)
        hipLaunchKernelGGL((k_mfma_tiles<E_LOCALMAX, KB, PF, OCC>), dim3(grid), dim3(NTHR), 0, s, a);
    else if (epi == E_TBE)
        hipLaunchKernelGGL((k_mfma_tiles<E_TBE, KB, PF, OCC>), dim3(grid), dim3(NTHR), 0, s, a);
    else if (epi == E_PAIRS)
        hipLaunchKernelGGL((k_mfma_tiles<E_PAIRS, KB, PF, OCC>), dim3(grid), dim3(NTHR), 0, s, a);
    else if (epi == E_ARGMAX)
        hipLaunchKernelGGL((k_mfma_tiles<E_ARGMAX, KB, PF, OCC>), dim3(grid), dim3(NTHR), 0, s, a);
    else
        throw Error(CASPHY_EINTERNAL, "MFMA tiles: unsupported epilogue");
}

void launch_wide(int epi, const MfmaArgs &a, hipStream_t s)
{
    const u32 grid = 8 * ((a.ntiles + 7) / 8);
    if (epi == E_MASK)
        hipLaunchKernelGGL(k_mfma_wide<E_MASK>, dim3(grid), dim3(256), 0, s, a);
    else if (epi == E_SCORES)
        hipLaunchKernelGGL(k_mfma_wide<E_SCORES>, dim3(grid), dim3(256), 0, s, a);
    else
        throw Error(CASPHY_EINTERNAL, "MFMA wide tiles: unsupported epilogue");
}

void launch_wide8(int epi, const MfmaArgs &a, hipStream_t s)
{
    const u32 grid = 8 * ((a.ntiles + 7) / 8);
    if (epi == E_MASK)
        hipLaunchKernelGGL(k_mfma_wide8<E_MASK>, dim3(grid), dim3(512), 0, s, a);
    else if (epi == E_SCORES)
        hipLaunchKernelGGL(k_mfma_wide8<E_SCORES>, dim3(grid), dim3(512), 0, s, a);
    else
        throw Error(CASPHY_EINTERNAL, "MFMA wide tiles: unsupported epilogue");
}

// Expands the packed operands to fp4 nibbles (stream-ordered scratch freed
// after the launch) and runs the fp4 tiles.
// mode 0: LDS staging through registers, 1: B operands from global, 2: LDS DMA,
// 3: LDS DMA with 256 x 256 tiles, 4: 256 x 256 tiles of 4 software-pipelined waves,
// 5: 256 x 128 tiles of 4 pipelined waves, two workgroups per CU, 6: 256 x 256
// tiles of 4 pipelined waves, one persistent workgroup per CU
void launch_fp4(int epi, const MfmaArgs &in, hipStream_t s, int mode)
{
    MfmaArgs a = in;
    const u32 stages = a.stages;
    DevBuf ea, eb;
    const u32 *EA = a.EA, *EB = a.EB;
    if (!EA) {
        ea.alloc(fp4_bytes(a.padA, stages));
        launch_expand_fp4(a.PA, a.padA, stages, ea.as<u32>(), s);
        EA = ea.as<u32>();
    }
    if (!EB) {
        if (a.PB == a.PA && a.padB == a.padA) {
            EB = EA;
        } else {
            eb.alloc(fp4_bytes(a.padB, stages));
            launch_expand_fp4(a.PB, a.padB, stages, eb.as<u32>(), s);
            EB = eb.as<u32>();
        }
    }
    a.PA = EA;
    a.PB = EB;
    const u32 grid = 8 * ((a.ntiles + 7) / 8);
    if ((mode == 4 || mode == 5) && (stages % 4 != 0 || stages < 4))
        throw Error(CASPHY_EINTERNAL, "MFMA fp4 tiles: the pipelined tiles need a multiple of 4 stages");
    if (mode == 6) {
        // one workgroup per CU (the XCD mapping wants a multiple of 8); the CU
        // count of each device is looked up once (any thread may race to it,
        // all store the same value)
        static std::atomic<int> cu_count[64];
        int dev = 0;
        CASPHY_HIP(hipGetDevice(&dev));
        int cus = dev >= 0 && dev < 64 ? cu_count[dev].load(std::memory_order_relaxed) : 0;
        if (!cus) {
            CASPHY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            if (dev >= 0 && dev < 64) cu_count[dev].store(cus, std::memory_order_relaxed);
        }
        const u32 per_xcd = std::max(1u, (u32)cus / 8), per = (a.ntiles + 7) / 8;
        const u32 pgrid = 8 * std::min(per_xcd, per);
        if (epi == E_MASK)
            hipLaunchKernelGGL(k_mfma_fp4p<E_MASK>, dim3(pgrid), dim3(256), 0, s, a);
        else if (epi == E_SCORES)
            hipLaunchKernelGGL(k_mfma_fp4p<E_SCORES>, dim3(pgrid), dim3(256), 0, s, a);
        else if (epi == E_LOCALMAX)
            hipLaunchKernelGGL(k_mfma_fp4p<E_LOCALMAX>, dim3(pgrid), dim3(256), 0, s, a);
        else if (epi == E_TBE)
            hipLaunchKernelGGL(k_mfma_fp4p<E_TBE>, dim3(pgrid), dim3(256), 0, s, a);
        else if (epi == E_PAIRS)
            hipLaunchKernelGGL(k_mfma_fp4p<E_PAIRS>, dim3(pgrid), dim3(256), 0, s, a);
        else if (epi == E_ARGMAX)
            hipLaunchKernelGGL(k_mfma_fp4p<E_ARGMAX>, dim3(pgrid), dim3(256), 0, s, a);
        else
            throw Error(CASPHY_EINTERNAL, "MFMA fp4 tiles: unsupported epilogue");
    } else if (mode == 5) {
        if (epi == E_MASK)
            hipLaunchKernelGGL(k_mfma_fp4y<E_MASK>, dim3(grid), dim3(256), 0, s, a);
        else if (epi == E_SCORES)
            hipLaunchKernelGGL(k_mfma_fp4y<E_SCORES>, dim3(grid), dim3(256), 0, s, a);
        else if (epi == E_LOCALMAX)
            hipLaunchKernelGGL(k_mfma_fp4y<E_LOCALMAX>, dim3(grid), dim3(256), 0, s, a);
        else
            throw Error(CASPHY_EINTERNAL, "MFMA fp4 tiles: unsupported epilogue");
    } else if (mode == 4) {
        if (epi == E_MASK)
            hipLaunchKernelGGL(k_mfma_fp4x<E_MASK>, dim3(grid), dim3(256), 0, s, a);
        else if (epi == E_SCORES)
            hipLaunchKernelGGL(k_mfma_fp4x<E_SCORES>, dim3(grid), dim3(256), 0, s, a);
        else if (epi == E_LOCALMAX)
            hipLaunchKernelGGL(k_mfma_fp4x<E_LOCALMAX>, dim3(grid), dim3(256), 0, s, a);
        else if (epi == E_TBE)
            hipLaunchKernelGGL(k_mfma_fp4x<E_TBE>, dim3(grid), dim3(256), 0, s, a);
        else if (epi == E_PAIRS)
            hipLaunchKernelGGL(k_mfma_fp4x<E_PAIRS>, dim3(grid), dim3(256), 0, s, a);
        else if (epi == E_ARGMAX)
            hipLaunchKernelGGL(k_mfma_fp4x<E_ARGMAX>, dim3(grid), dim3(256), 0, s, a);
        else
            throw Error(CASPHY_EINTERNAL, "MFMA fp4 tiles: unsupported epilogue");
    } else if (mode == 3) {
        if (epi == E_MASK)
            hipLaunchKernelGGL(k_mfma_fp4w<E_MASK>, dim3(grid), dim3(512), 0, s, a);
        else if (epi == E_SCORES)
            hipLaunchKernelGGL(k_mfma_fp4w<E_SCORES>, dim3(grid), dim3(512), 0, s, a);
        else if (epi == E_LOCALMAX)
            hipLaunchKernelGGL(k_mfma_fp4w<E_LOCALMAX>, dim3(grid), dim3(512), 0, s, a);
        else if (epi == E_TBE)
            hipLaunchKernelGGL(k_mfma_fp4w<E_TBE>, dim3(grid), dim3(512), 0, s, a);
        else if (epi == E_PAIRS)
            hipLaunchKernelGGL(k_mfma_fp4w<E_PAIRS>, dim3(grid), dim3(512), 0, s, a);
        else if (epi == E_ARGMAX)
            hipLaunchKernelGGL(k_mfma_fp4w<E_ARGMAX>, dim3(grid), dim3(512), 0, s, a);
        else
            throw Error(CASPHY_EINTERNAL, "MFMA fp4 tiles: unsupported epilogue");
    } else if (mode == 2) {
        if (epi == E_MASK)
            hipLaunchKernelGGL(k_mfma_fp4d<E_MASK>, dim3(grid), dim3(NTHR), 0, s, a);
        else if (epi == E_SCORES)
            hipLaunchKernelGGL(k_mfma_fp4d<E_SCORES>, dim3(grid), dim3(NTHR), 0, s, a);
        else if (epi == E_LOCALMAX)
            hipLaunchKernelGGL(k_mfma_fp4d<E_LOCALMAX>, dim3(grid), dim3(NTHR), 0, s, a);
        else if (epi == E_TBE)
            hipLaunchKernelGGL(k_mfma_fp4d<E_TBE>, dim3(grid), dim3(NTHR), 0, s, a);
        else if (epi == E_PAIRS)
            hipLaunchKernelGGL(k_mfma_fp4d<E_PAIRS>, dim3(grid), dim3(NTHR), 0, s, a);
        else if (epi == E_ARGMAX)
            hipLaunchKernelGGL(k_mfma_fp4d<E_ARGMAX>, dim3(grid), dim3(NTHR), 0, s, a);
        else
            throw Error(CASPHY_EINTERNAL, "MFMA fp4 tiles: unsupported epilogue");
    } else if (mode == 1) {
        if (epi == E_MASK)
            hipLaunchKernelGGL(k_mfma_fp4b<E_MASK>, dim3(grid), dim3(NTHR), 0, s, a);
        else if (epi == E_SCORES)
            hipLaunchKernelGGL(k_mfma_fp4b<E_SCORES>, dim3(grid), dim3(NTHR), 0, s, a);
        else if (epi == E_LOCALMAX)
            hipLaunchKernelGGL(k_mfma_fp4b<E_LOCALMAX>, dim3(grid), dim3(NTHR), 0, s, a);
        else if (epi == E_TBE)
            hipLaunchKernelGGL(k_mfma_fp4b<E_TBE>, dim3(grid), dim3(NTHR), 0, s, a);
        else if (epi == E_PAIRS)
            hipLaunchKernelGGL(k_mfma_fp4b<E_PAIRS>, dim3(grid), dim3(NTHR), 0, s, a);
        else if (epi == E_ARGMAX)
            hipLaunchKernelGGL(k_mfma_fp4b<E_ARGMAX>, dim3(grid), dim3(NTHR), 0, s, a);
        else
            throw Error(CASPHY_EINTERNAL, "MFMA fp4 tiles: unsupported epilogue");
    } else if (epi == E_MASK) {
        hipLaunchKernelGGL(k_mfma_fp4<E_MASK>, dim3(grid), dim3(NTHR), 0, s, a);
    } else if (epi == E_SCORES) {
        hipLaunchKernelGGL(k_mfma_fp4<E_SCORES>, dim3(grid), dim3(NTHR), 0, s, a);
    } else if (epi == E_LOCALMAX) {
        hipLaunchKernelGGL(k_mfma_fp4<E_LOCALMAX>, dim3(grid), dim3(NTHR), 0, s, a);
    } else if (epi == E_TBE) {
        hipLaunchKernelGGL(k_mfma_fp4<E_TBE>, dim3(grid), dim3(NTHR), 0, s, a);
    } else if (epi == E_PAIRS) {
        hipLaunchKernelGGL(k_mfma_fp4<E_PAIRS>, dim3(grid), dim3(NTHR), 0, s, a);
    } else if (epi == E_ARGMAX) {
        hipLaunchKernelGGL(k_mfma_fp4<E_ARGMAX>, dim3(grid), dim3(NTHR), 0, s, a);
    } else {
        throw Error(CASPHY_EINTERNAL, "MFMA fp4 tiles: unsupported epilogue");
    }
}

void launch_reg(int epi, const MfmaArgs &a, hipStream_t s)
{
    const u32 grid = 8 * ((a.ntiles + 7) / 8);
    if (epi == E_MASK)
        hipLaunchKernelGGL(k_mfma_reg<E_MASK>, dim3(grid), dim3(256), 0, s, a);
    else if (epi == E_SCORES)
        hipLaunchKernelGGL(k_mfma_reg<E_SCORES>, dim3(grid), dim3(256), 0, s, a);
    else if (epi == E_LOCALMAX)
        hipLaunchKernelGGL(k_mfma_reg<E_LOCALMAX>, dim3(grid), dim3(256), 0, s, a);
    else if (epi == E_TBE)
        hipLaunchKernelGGL(k_mfma_reg<E_TBE>, dim3(grid), dim3(256), 0, s, a);
    else if (epi == E_PAIRS)
        hipLaunchKernelGGL(k_mfma_reg<E_PAIRS>, dim3(grid), dim3(256), 0, s, a);
    else if (epi == E_ARGMAX)
        hipLaunchKernelGGL(k_mfma_reg<E_ARGMAX>, dim3(grid), dim3(256), 0, s, a);
    else
        throw Error(CASPHY_EINTERNAL, "MFMA tiles: unsupported epilogue");
}

namespace {

// Variants: LDS-staged with stage depth x prefetch, or register-resident ("reg").
const char *const kVariants[] = {"128x1", "128x2", "128x4", "64x2", "64x4", "64x8", "reg", "wide", "wide8", "fp4",
                                 "fp4b", "fp4d", "fp4w", "fp4x", "fp4y", "fp4p"};
constexpr int kNumVariants = 16;
constexpr int kDefaultVariant = 3;
constexpr int kCandidates[] = {3, 0, 4, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15};  // what the autotuner times
constexpr int kWide = 7, kWide8 = 8;  // radius-mask and score epilogues only
constexpr int kFp4 = 9, kFp4b = 10, kFp4d = 11, kFp4w = 12, kFp4x = 13, kFp4y = 14;  // fp4y: mask, scores, localmax
constexpr int kFp4p = 15;
// super-tile edge for 128 x 128 tiles: 8 (8 x 8 tiles = 16 row blocks of 128);
// CASPHY_MFMA_SUPER overrides, 0 keeps the plain column order
u32 super_edge()
{
    const char *e = std::getenv("CASPHY_MFMA_SUPER");
    return e ? (u32)std::atoi(e) : 8u;
}

// super-tile edge of the persistent 256 x 256 tiles; CASPHY_FP4P_SUPER
// overrides, 0 keeps the plain column order
u32 fp4p_super_edge()
{
    const char *e = std::getenv("CASPHY_FP4P_SUPER");
    return e ? (u32)std::atoi(e) : 0u;
}

bool variant_supports(int v, int epi)
{
    if (v == kFp4y) return epi == E_MASK || epi == E_SCORES || epi == E_LOCALMAX;
    return v < kWide || v >= kFp4 || epi == E_MASK || epi == E_SCORES;
}

int variant_from_env()
{
    const char *e = std::getenv("CASPHY_MFMA_VARIANT");
    if (e)
        for (int k = 0; k < kNumVariants; ++k)
            if (!std::strcmp(e, kVariants[k])) return k;
    return -1;
}

void launch_one(int v, int epi, const MfmaArgs &in, hipStream_t s)
{
    MfmaArgs a = in;
    if (!variant_supports(v, epi)) v = kDefaultVariant;
    // the persistent tiles keep 32-bit DMA offsets (below 2^26 padded rows) and
    // prefetch four stages ahead
    if (v == kFp4p && (a.padA >= (1u << 26) || a.padB >= (1u << 26) || a.stages < 4)) v = kFp4;
    const u32 T = v == 6 ? RT
                         : (v == kWide || v == kWide8 || v == kFp4w || v == kFp4x || v == kFp4y || v == kFp4p) ? WT
                                                                                                               : TILE;
    const u32 ntA = ceil_div(a.nA, T);
    a.ntA = ntA;
    a.ntB = ceil_div(a.nB, v == kFp4y ? (u32)TILE : T);  // fp4y: 256 A rows x 128 B rows
    if (ntA == 0 || a.ntB == 0) return;
    if (v == kFp4y) {
        // triangle: the two 128-row halves of every 256-row block pair I <= J
        a.sup = 0;
        a.ntiles = a.triangle ? (u32)((u64)ntA * (ntA + 1)) : ntA * a.ntB;
        launch_fp4(epi, a, s, 5);
        CASPHY_HIP(hipGetLastError());
        return;
    }
    // 128 x 128 tiles run in 8 x 8 super-tiles (launch slots without a tile exit)
    a.sup = T == TILE ? super_edge() : v == kFp4p ? fp4p_super_edge() : 0;
    if (a.sup) {
        const u64 nSA = ceil_div(ntA, a.sup), nSB = ceil_div(a.ntB, a.sup);
        const u64 nsup = a.triangle ? nSA * (nSA + 1) / 2 : nSA * nSB;
        a.ntiles = (u32)(nsup * a.sup * a.sup);
    } else {
        a.ntiles = a.triangle ? (u32)((u64)ntA * (ntA + 1) / 2) : ntA * a.ntB;
    }
    switch (v) {
    case 0: launch_variant<128, 1, 2>(epi, a, s); break;
    case 1: launch_variant<128, 2, 2>(epi, a, s); break;
    case 2: launch_variant<128, 4, 2>(epi, a, s); break;
    case 3: launch_variant<64, 2, 4>(epi, a, s); break;
    case 4: launch_variant<64, 4, 4>(epi, a, s); break;
    case 5: launch_variant<64, 8, 4>(epi, a, s); break;
    case 6: launch_reg(epi, a, s); break;
    case kWide: launch_wide(epi, a, s); break;
    case kWide8: launch_wide8(epi, a, s); break;
    case kFp4: launch_fp4(epi, a, s, 0); break;
    case kFp4b: launch_fp4(epi, a, s, 1); break;
    case kFp4d: launch_fp4(epi, a, s, 2); break;
    case kFp4w: launch_fp4(epi, a, s, 3); break;
    case kFp4x: launch_fp4(epi, a, s, 4); break;
    case kFp4y: launch_fp4(epi, a, s, 5); break;
    case kFp4p: launch_fp4(epi, a, s, 6); break;
    default: throw Error(CASPHY_EINTERNAL, "MFMA tiles: unknown variant");
    }
    CASPHY_HIP(hipGetLastError());
}

// Autotuner: the first launch of each shape class (epilogue, triangle,
// log2 of both row counts, sites) runs every candidate variant twice on the
// real data -- all epilogues are idempotent (mask / score stores, atomic max /
// min of identical values) -- times the second run with events and keeps the
// fastest.
struct TuneKey {
    int epi, tri;
    u32 la, lb, stages;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(epi, tri, la, lb, stages) < std::tie(o.epi, o.tri, o.la, o.lb, o.stages);
    }
};
struct TuneEntry {
    int variant;
    float ms[kNumVariants];
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneEntry> g_tune;

u32 log2_ceil(u32 x)
{
    u32 l = 0;
    while ((1u << l) < x && l < 31) ++l;
    return l;
}

}  // namespace

void launch_mfma_tiles(int epi, const MfmaArgs &a, hipStream_t s)
{
    // the fp4 tiles sum in fp32, exact below 2^24: wider rows (every partial
    // sum is at most the number of sites) take the int32 sums of the int8 tiles
    const bool fp4_exact = (u64)a.stages * 128 < (1ull << 24);
    int forced = variant_from_env();
    if (forced >= kFp4 && !fp4_exact) forced = kDefaultVariant;
    if (forced >= 0) return launch_one(forced, epi, a, s);
    if (!fp4_exact) return launch_one(kDefaultVariant, epi, a, s);
    const char *at = std::getenv("CASPHY_AUTOTUNE");
    if (at && at[0] == '0') return launch_one(kDefaultVariant, epi, a, s);
    const TuneKey key{epi, a.triangle, log2_ceil(a.nA), log2_ceil(a.nB), a.stages};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) return launch_one(it->second.variant, epi, a, s);
    }
    // under stream capture nothing can be timed: the default variant
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    CASPHY_HIP(hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone) return launch_one(kDefaultVariant, epi, a, s);
    // a radius mask whose threshold lives in device memory may be a launch that
    // scores nothing (threshold INT32_MAX): the candidates are timed with a
    // host threshold instead, and the real launch follows with the choice
    MfmaArgs t = a;
    const bool relaunch = epi == E_MASK && a.thr_dev;
    if (relaunch) {
        t.thr_dev = nullptr;
        t.thr = 1;
    }
    TuneEntry e{};
    for (float &m : e.ms) m = -1.0f;
    hipEvent_t ev[2];
    CASPHY_HIP(hipEventCreate(&ev[0]));
    CASPHY_HIP(hipEventCreate(&ev[1]));
    float best = 1e30f;
    e.variant = kDefaultVariant;
    for (int v : kCandidates) {
        if (!variant_supports(v, epi)) continue;
        launch_one(v, epi, t, s);  // untimed: first-use costs (code load, scratch pool growth)
        CASPHY_HIP(hipEventRecord(ev[0], s));
        launch_one(v, epi, t, s);
        CASPHY_HIP(hipEventRecord(ev[1], s));
        CASPHY_HIP(hipEventSynchronize(ev[1]));
        float ms = 0;
        CASPHY_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
        e.ms[v] = ms;
        if (ms < best) { best = ms; e.variant = v; }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (relaunch) launch_one(e.variant, epi, a, s);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.emplace(key, e);
}

namespace {
std::vector<std::string> g_notes;  // guarded by g_tune_mu
}

void add_tuning_note(const std::string &json_object)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_notes.push_back(json_object);
}

void launch_mfma_tiles_untuned(int epi, const MfmaArgs &a, hipStream_t s)
{
    const bool fp4_exact = (u64)a.stages * 128 < (1ull << 24);
    int v = variant_from_env();
    if (v < 0) {
        const TuneKey key{epi, a.triangle, log2_ceil(a.nA), log2_ceil(a.nB), a.stages};
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        v = it != g_tune.end() ? it->second.variant : kFp4;
    }
    if (v >= kFp4 && (!fp4_exact || !a.EA || !a.EB)) v = kDefaultVariant;
    launch_one(v, epi, a, s);
}

namespace {

// ctab[x] = #{k : radii[k] <= radii[0] + x} for x < span = radii[nrad-1] -
// radii[0] + 1 (at most COUNT_SPAN_MAX), cneed[k] = #{j : radii[j] <=
// radii[k]}, cpar = (radii[0], span).  Radii ascending.
__global__ void k_count_table(const int32_t *radii, u32 nrad, unsigned char *ctab, unsigned char *cneed,
                              int32_t *cpar)
{
    const int32_t rmin = radii[0];
    const u32 span = (u32)min((int64_t)radii[nrad - 1] - rmin + 1, (int64_t)COUNT_SPAN_MAX);
    for (u32 x = threadIdx.x; x < span; x += blockDim.x) {
        const int64_t sc = (int64_t)rmin + x;
        u32 c = 0;
        for (u32 k = 0; k < nrad; ++k) c += radii[k] <= sc;
        ctab[x] = (unsigned char)c;
    }
    for (u32 k = threadIdx.x; k < nrad; k += blockDim.x) {
        u32 c = 0;
        for (u32 j = 0; j < nrad; ++j) c += radii[j] <= radii[k];
        cneed[k] = (unsigned char)c;
    }
    if (threadIdx.x == 0) {
        cpar[0] = rmin;
        cpar[1] = (int32_t)span;
    }
}

// counts[k * R + q] = sum over the ntB column tiles of cpart[(J * nrad + k) * padA + q];
// one wave per radius and 4 consecutive rows, its lanes over the column tiles
// (independent loads), then a wave sum.
__global__ __launch_bounds__(256) void k_count_sum(const unsigned char *cpart, u32 ntB, u32 nrad, u32 padA, u32 R,
                                                   u32 *counts)
{
    const u32 lane = threadIdx.x & 63;
    const u32 q4 = padA / 4;
    const u64 items = (u64)nrad * q4;
    const size_t stride = (size_t)nrad * padA / 4;
    for (u64 it = blockIdx.x * 4ull + (threadIdx.x >> 6); it < items; it += gridDim.x * 4ull) {
        const u32 k = (u32)(it / q4), q = 4 * (u32)(it - (u64)k * q4);
        if (q >= R) continue;  // wave-uniform
        const u32 *p = reinterpret_cast<const u32 *>(cpart + (size_t)k * padA + q);
        u32 c0 = 0, c1 = 0, c2 = 0, c3 = 0;
        for (u32 J = lane; J < ntB; J += 64) {
            const u32 w = p[(size_t)J * stride];
            c0 += w & 0xFFu; c1 += (w >> 8) & 0xFFu; c2 += (w >> 16) & 0xFFu; c3 += w >> 24;
        }
        for (int o = 32; o > 0; o >>= 1) {
            c0 += __shfl_xor(c0, o); c1 += __shfl_xor(c1, o); c2 += __shfl_xor(c2, o); c3 += __shfl_xor(c3, o);
        }
        if (lane == 0) {
            u32 *out = counts + (size_t)k * R + q;
            out[0] = c0;
            if (q + 1 < R) out[1] = c1;
            if (q + 2 < R) out[2] = c2;
            if (q + 3 < R) out[3] = c3;
        }
    }
}

}  // namespace

bool fused_counts_ok(const MfmaArgs &a, u32 nradii)
{
    const char *e = std::getenv("CASPHY_FUSED_COUNTS");
    if (e && e[0] == '0') return false;
    const int v = variant_from_env();
    if (v >= 0 && v != kFp4) return false;  // a forced variant keeps the score matrix
    return a.EA && a.EB && (u64)a.stages * 128 < (1ull << 24) && nradii >= 1 && nradii <= MAX_ROUND_RADII_FUSED;
}

void launch_count_table(const int32_t *radii, u32 nrad, unsigned char *ctab, unsigned char *cneed, int32_t *cpar,
                        hipStream_t s)
{
    hipLaunchKernelGGL(k_count_table, dim3(1), dim3(256), 0, s, radii, nrad, ctab, cneed, cpar);
    CASPHY_HIP(hipGetLastError());
}

void launch_mfma_counts(const MfmaArgs &in, u32 *counts, u32 R, hipStream_t s)
{
    MfmaArgs a = in;
    CASPHY_REQUIRE(a.EA && a.EB && a.ctab && a.cneed && a.cpar && a.cpart && a.nrad >= 1 &&
                       a.nrad <= MAX_ROUND_RADII_FUSED && a.triangle == 0 && a.padA % 4 == 0,
                   CASPHY_EINTERNAL, "fused tryRadius counts: bad arguments");
    a.ntA = ceil_div(a.nA, TILE);
    a.ntB = ceil_div(a.nB, TILE);
    if (a.ntA == 0 || a.ntB == 0) return;
    a.sup = super_edge();
    if (a.sup) {
        const u64 nSA = ceil_div(a.ntA, a.sup), nSB = ceil_div(a.ntB, a.sup);
        a.ntiles = (u32)(nSA * nSB * a.sup * a.sup);
    } else {
        a.ntiles = a.ntA * a.ntB;
    }
    a.PA = a.EA;
    a.PB = a.EB;
    const u32 grid = 8 * ((a.ntiles + 7) / 8);
    hipLaunchKernelGGL(k_mfma_fp4<E_COUNTS>, dim3(grid), dim3(NTHR), 0, s, a);
    CASPHY_HIP(hipGetLastError());
    const u64 items = (u64)a.nrad * (a.padA / 4);
    hipLaunchKernelGGL(k_count_sum, dim3((u32)std::min<u64>((items + 3) / 4, 16384)), dim3(256), 0, s, a.cpart,
                       a.ntB, a.nrad, a.padA, R, counts);
    CASPHY_HIP(hipGetLastError());
}

void launch_mfma_groups(int epi, const MfmaArgs &in, u32 nslots, hipStream_t s)
{
    CASPHY_REQUIRE(in.glist && in.gdesc && (epi == E_MASK || epi == E_SCORES), CASPHY_EINTERNAL,
                   "MFMA group launch: bad arguments");
    if (!nslots) return;
    MfmaArgs a = in;
    a.ntiles = nslots;
    a.sup = 0;
    const bool fp4_exact = (u64)a.stages * 128 < (1ull << 24);
    int v = variant_from_env();
    if (v < 0 || v == 6 || v == kWide || v == kWide8 || v == kFp4w) v = kFp4;  // 128 x 128 tiles only
    if (v >= kFp4 && (!fp4_exact || !a.EA || !a.EB)) v = kDefaultVariant;
    switch (v) {
    case 0: launch_variant<128, 1, 2>(epi, a, s); break;
    case 1: launch_variant<128, 2, 2>(epi, a, s); break;
    case 2: launch_variant<128, 4, 2>(epi, a, s); break;
    case 3: launch_variant<64, 2, 4>(epi, a, s); break;
    case 4: launch_variant<64, 4, 4>(epi, a, s); break;
    case 5: launch_variant<64, 8, 4>(epi, a, s); break;
    case kFp4: launch_fp4(epi, a, s, 0); break;
    case kFp4b: launch_fp4(epi, a, s, 1); break;
    default: launch_fp4(epi, a, s, 2); break;
    }
    CASPHY_HIP(hipGetLastError());
}

std::string mfma_tuning_report()
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::string out = "[";
    for (auto &note : g_notes) {
        if (out.size() > 1) out += ",";
        out += note;
    }
    for (auto &kv : g_tune) {
        if (out.size() > 1) out += ",";
        out += "{\"epilogue\":" + std::to_string(kv.first.epi) + ",\"triangle\":" + std::to_string(kv.first.tri) +
               ",\"log2_rows_a\":" + std::to_string(kv.first.la) + ",\"log2_rows_b\":" + std::to_string(kv.first.lb) +
               ",\"sites\":" + std::to_string(kv.first.stages * 128) + ",\"chosen\":\"" + kVariants[kv.second.variant] +
               "\",\"ms\":{";
        bool first = true;
        for (int v : kCandidates) {
            if (kv.second.ms[v] < 0) continue;
            if (!first) out += ",";
            first = false;
            out += "\"" + std::string(kVariants[v]) + "\":" + std::to_string(kv.second.ms[v]);
        }
        out += "}}";
    }
    return out + "]";
}

}  // namespace casphy
