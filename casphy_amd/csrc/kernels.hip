// kernels.hip -- layout conversion and radius-list compaction kernels.
//
// All of these are HBM-streaming kernels: coalesced 4/16-byte accesses, one
// wave per row where a row is scanned, no atomics.
#include <algorithm>

#include "kernels.hpp"

namespace casphy {

namespace {

__global__ void k_transpose_wordmajor(const u32 *in, u32 W, u32 n, u32 *out, u32 ldw)
{
    // 32 x 32 tile through LDS: reads along n (contiguous in .bmat), writes along words
    __shared__ u32 t[32][33];
    const u32 w0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
    const u32 tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 8 rows per pass
    for (u32 y = ty; y < 32; y += 8) {
        const u32 w = w0 + y, c = c0 + tx;
        t[y][tx] = (w < W && c < n) ? in[(size_t)w * n + c] : 0u;
    }
    __syncthreads();
    for (u32 y = ty; y < 32; y += 8) {
        const u32 c = c0 + y, w = w0 + tx;
        if (c < n && w < ldw) out[(size_t)c * ldw + w] = t[tx][y];
    }
}

__global__ void k_pad_rows(const u32 *in, u32 W, u32 n, u32 *out, u32 ldw)
{
    const u64 total = (u64)n * ldw;
    for (u64 x = blockIdx.x * (u64)blockDim.x + threadIdx.x; x < total; x += (u64)gridDim.x * blockDim.x) {
        const u32 r = (u32)(x / ldw), w = (u32)(x - (u64)r * ldw);
        out[x] = w < W ? in[(size_t)r * W + w] : 0u;
    }
}

// A nibble with two or more bits set is an IUPAC ambiguity code.
__device__ __forceinline__ bool has_multibit_nibble(u32 v)
{
    const u32 nib = v & 0x0F0F0F0Fu;
    // x & (x - 1) per byte without borrow across bytes: nibbles are < 16
    const u32 m = nib & ((nib | 0x10101010u) - 0x01010101u) & 0x0F0F0F0Fu;
    return m != 0;
}

__global__ void k_dna_detect(const u32 *raw, u64 nwords, u32 *flag)
{
    bool any = false;
    for (u64 x = blockIdx.x * (u64)blockDim.x + threadIdx.x; x < nwords; x += (u64)gridDim.x * blockDim.x)
        any |= has_multibit_nibble(raw[x]);
    if (__any(any) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

// One thread per (row, 32-position group): reads 8 raw words.
// Nucleotide bit -> 2-bit code: A(1)->0, T(2)->1, C(4)->2, G(8)->3.
__global__ void k_dna_pack2(const u32 *raw, u32 n, u32 nwRaw, u32 *out, u32 ldw)
{
    const u32 groups = ldw / 3;
    const u64 total = (u64)n * groups;
    for (u64 x = blockIdx.x * (u64)blockDim.x + threadIdx.x; x < total; x += (u64)gridDim.x * blockDim.x) {
        const u32 r = (u32)(x / groups), g = (u32)(x - (u64)r * groups);
        u32 hi = 0, lo = 0, m = 0;
        for (u32 q = 0; q < 8; ++q) {
            const u32 w = g * 8 + q;
            const u32 v = w < nwRaw ? raw[(size_t)r * nwRaw + w] : 0u;
            for (u32 b = 0; b < 4; ++b) {
                const u32 nib = (v >> (8 * b)) & 0xFu;
                const u32 pos = q * 4 + b;
                if (nib) {
                    const u32 code = (nib & 0xCu ? 2u : 0u) | (nib & 0xAu ? 1u : 0u);
                    hi |= (code >> 1) << pos;
                    lo |= (code & 1u) << pos;
                    m |= 1u << pos;
                }
            }
        }
        u32 *o = out + (size_t)r * ldw + 3 * g;
        o[0] = hi; o[1] = lo; o[2] = m;
    }
}

__global__ void k_dna_pack_nib(const u32 *raw, u32 n, u32 nwRaw, u32 *out, u32 ldw)
{
    const u64 total = (u64)n * ldw;
    for (u64 x = blockIdx.x * (u64)blockDim.x + threadIdx.x; x < total; x += (u64)gridDim.x * blockDim.x) {
        const u32 r = (u32)(x / ldw), w = (u32)(x - (u64)r * ldw);
        u32 v = 0;
        for (u32 h = 0; h < 2; ++h) {
            const u32 rw = 2 * w + h;
            const u32 src = rw < nwRaw ? raw[(size_t)r * nwRaw + rw] : 0u;
            for (u32 b = 0; b < 4; ++b) v |= ((src >> (8 * b)) & 0xFu) << (16 * h + 4 * b);
        }
        out[x] = v;
    }
}

// One wave per row; rows starting on a 16-byte boundary are read as 16-byte
// words, four loads in flight per lane.
// A device threshold of INT32_MAX (skip) means the mask launch scored nothing:
// every count is 0 and the mask is not read.
// With map.moff set, row r is the row of map (its own offset, width and group).
__global__ void k_row_popc(const u32 *mask, u32 n, u32 mstride, u32 nwords, u32 *counts, const int32_t *skip,
                           RowMap map)
{
    const u32 lane = threadIdx.x & 63;
    const u32 row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= n) return;
    const int32_t *sk = map.moff ? map.gthr + 2 * map.group[row] : skip;
    if (sk && *sk == 0x7FFFFFFF) {
        if (lane == 0) counts[row] = 0;
        return;
    }
    const u32 *m = map.moff ? mask + map.moff[row] : mask + (size_t)row * mstride;
    const u32 words = map.moff ? map.words[row] : nwords;
    u32 c = 0, tail = 0;
    if ((reinterpret_cast<uintptr_t>(m) & 15) == 0) {
        const uint4 *m4 = reinterpret_cast<const uint4 *>(m);
        const u32 full = words / 4;
        for (u32 w0 = 0; w0 < full; w0 += 256) {
            uint4 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const u32 w = w0 + 64 * u + lane;
                v[u] = w < full ? m4[w] : make_uint4(0, 0, 0, 0);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                c += (u32)(__builtin_popcount(v[u].x) + __builtin_popcount(v[u].y) + __builtin_popcount(v[u].z) +
                           __builtin_popcount(v[u].w));
        }
        tail = full * 4;
    }
    for (u32 w = tail + lane; w < words; w += 64) c += (u32)__builtin_popcount(m[w]);
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane == 0) counts[row] = c;
}

// Single workgroup of 1024 threads over blocks of 4,096 counts: thread t
// loads counts 4t .. 4t+3 of the block, wave scans (shuffles) and a scan of the
// 16 wave totals give each thread its exclusive offset, plus the running total
// of the earlier blocks.
__global__ __launch_bounds__(1024) void k_scan_counts(const u32 *counts, u32 n, u64 *offsets)
{
    __shared__ u64 wsum[16];
    const u32 t = threadIdx.x, lane = t & 63, wv = t >> 6;
    u64 carry = 0;
    for (u32 base = 0; base < n; base += 4096) {
        const u32 i0 = base + 4 * t;
        u32 v[4];
        if (i0 + 3 < n && (reinterpret_cast<uintptr_t>(counts + i0) & 15) == 0) {
            const uint4 x = *reinterpret_cast<const uint4 *>(counts + i0);
            v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = i0 + u < n ? counts[i0 + u] : 0u;
        }
        const u64 mine = (u64)v[0] + v[1] + v[2] + v[3];
        u64 x = mine;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const u64 y = __shfl_up(x, o);
            if (lane >= (u32)o) x += y;
        }
        if (lane == 63) wsum[wv] = x;
        __syncthreads();
        if (wv == 0) {
            u64 w = lane < 16 ? wsum[lane] : 0;
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                const u64 y = __shfl_up(w, o);
                if (lane >= (u32)o) w += y;
            }
            if (lane < 16) wsum[lane] = w;
        }
        __syncthreads();
        u64 run = carry + (wv ? wsum[wv - 1] : 0) + x - mine;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i0 + u < n) {
                offsets[i0 + u] = run;
                run += v[u];
            }
        carry += wsum[15];
        __syncthreads();  // wsum is rewritten by the next block
    }
    if (t == 0) offsets[n] = carry;
}

// One wave per row, four rows per workgroup.  Per chunk of 64 mask words: a
// chunk of 64 full words is the run of ids 32 w0 .. 32 w0 + 2047 and goes
// straight to HBM as 16-byte stores (a few single ids align the run); any
// other chunk takes a wave prefix sum of the word popcounts that places every
// lane's ids in an LDS staging buffer at the position congruent, modulo 4, to
// their place in the list, so the wave streams the chunk's ids to HBM as
// aligned 16-byte stores (single ids at the two ends).  Rows with an empty
// list are not read.
// The mask words are loaded CSEG chunks at a time into registers: on this
// target the wait for a load also waits for every store issued before it, so
// loading a chunk just ahead of its use would drain the id stores once per
// chunk; a segment drains them once per CSEG chunks.  Measured on 2.8e9 ids
// from a 100,000-row mask: 2.52 ms, against 2.98 ms with a one-chunk prefetch
// and non-temporal stores (which lose to plain stores once the drains are
// rare).
typedef u32 v4u __attribute__((ext_vector_type(4)));

constexpr int CWAVES = 4;   // rows (waves) per workgroup
constexpr int CSEG = 25;    // 64-word chunks per register segment

// The list of one row; CAP: some ids fall at or past `cap` and are dropped.
template <bool CAP>
__device__ __forceinline__ void compact_row(const u32 *m, u32 words, u64 base, u32 *idx, u64 cap, u32 *buf, u32 lane)
{
    for (u32 s0 = 0; s0 < words; s0 += 64 * CSEG) {
        u32 seg[CSEG];
#pragma unroll
        for (int k = 0; k < CSEG; ++k) {
            const u32 w = s0 + 64 * k + lane;
            seg[k] = w < words ? m[w] : 0u;
        }
#pragma unroll
        for (int k = 0; k < CSEG; ++k) {
            const u32 w0 = s0 + 64 * k, w = w0 + lane;
            if (w0 >= words) break;  // wave-uniform
            u32 v = seg[k];
            if (w0 + 64 <= words && __all(v == 0xFFFFFFFFu)) {  // wave-uniform
                const u32 id0 = w0 * 32;
                const u32 head = (u32)((4 - (base & 3)) & 3);
                if (lane < head && (!CAP || base + lane < cap)) idx[base + lane] = id0 + lane;
                const u64 b4 = base + head;  // multiple of 4: 16-byte aligned
                const u32 nq = (2048 - head) >> 2;
                v4u *dst = reinterpret_cast<v4u *>(idx + b4);
                for (u32 q = lane; q < nq; q += 64) {
                    const u32 kq = head + 4 * q;
                    if (!CAP || b4 + 4 * q + 4 <= cap) {
                        dst[q] = v4u{id0 + kq, id0 + kq + 1, id0 + kq + 2, id0 + kq + 3};
                    } else {
                        for (u32 e = 0; e < 4; ++e)
                            if (b4 + 4 * q + e < cap) idx[b4 + 4 * q + e] = id0 + kq + e;
                    }
                }
                const u32 done = head + 4 * nq;
                if (lane < 2048 - done && (!CAP || base + done + lane < cap)) idx[base + done + lane] = id0 + done + lane;
                base += 2048;
                continue;
            }
            const u32 c = (u32)__builtin_popcount(v);
            const u32 inc = wave_inclusive_scan(c);
            const u32 total = __builtin_amdgcn_readlane(inc, 63);
            if (total == 0) continue;  // wave-uniform
            const u32 sh = (u32)(base & 3);  // list position base + j sits at buf[sh + j]
            u32 p = sh + inc - c;
            while (v) {
                const u32 b = (u32)__builtin_ctz(v);
                buf[p++] = w * 32 + b;
                v &= v - 1;
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            // 16-byte groups g of the list positions base - sh + 4g .. + 3 (aligned in
            // HBM and in LDS); the first and last group may hold positions of other
            // chunks, which are left alone
            const u64 a0 = base - sh;
            const u32 ng = (sh + total + 3) >> 2;
            for (u32 g = lane; g < ng; g += 64) {
                const u32 lo = 4 * g;
                if (lo >= sh && lo + 4 <= sh + total && (!CAP || a0 + lo + 4 <= cap)) {
                    *reinterpret_cast<v4u *>(idx + a0 + lo) = *reinterpret_cast<const v4u *>(buf + lo);
                } else {
                    for (u32 e = 0; e < 4; ++e) {
                        const u32 kk = lo + e;
                        if (kk >= sh && kk < sh + total && (!CAP || a0 + kk < cap)) idx[a0 + kk] = buf[kk];
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            base += total;
        }
    }
}

__global__ __launch_bounds__(64 * CWAVES) void k_compact(const u32 *mask, u32 n, u32 mstride, u32 nwords,
                                                        const u64 *offsets, u32 *idx, u64 cap, RowMap map)
{
    __shared__ __attribute__((aligned(16))) u32 stage[CWAVES][64 * 32 + 4];
    const u32 lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const u32 row = blockIdx.x * CWAVES + wv;
    if (row >= n) return;
    const u64 base = offsets[row];
    if (offsets[row + 1] == base) return;  // empty list: the mask row is not read
    const u32 *m = map.moff ? mask + map.moff[row] : mask + (size_t)row * mstride;
    const u32 words = map.moff ? map.words[row] : nwords;
    if (offsets[row + 1] > cap) compact_row<true>(m, words, base, idx, cap, stage[wv], lane);
    else compact_row<false>(m, words, base, idx, cap, stage[wv], lane);
}

// tryRadius counts from a query x cell score matrix: one workgroup per query.
//   * radii in 0 .. COUNT_LUT - 2 (the usual case): a table in LDS maps a score
//     x to the one-hot byte field idx(x) - 1, idx(x) = #{k : radii[k] <= x}, so
//     a score costs a clamp, one LDS read and four adds into four words of
//     byte counters (flushed every 255 scores); radius k then counts the
//     scores with idx >= #{j : radii[j] <= radii[k]};
//   * other radii: one compare per radius and score.
// Rows starting on a 16-byte boundary are read as 16-byte words, four loads in
// flight per lane.  The per-thread counters are summed through LDS.
constexpr u32 COUNT_LUT = 1024;  // table entries: scores -1 .. COUNT_LUT - 2

__global__ __launch_bounds__(256) void k_count_radii(const int32_t *S, u32 R, u32 nall, u64 ldS, const int32_t *radii,
                                                     u32 nradii, u32 *counts, const u64 *rowoff, const u32 *rowlen)
{
    __shared__ uint4 lds4[1088];  // the table, then 256 x 17 u32 for the reduction
    __shared__ u32 tot[MAX_RADII_K];
    const u32 q = blockIdx.x, tid = threadIdx.x;
    if (q >= R) return;
    int32_t rr[MAX_RADII_K];
    int32_t rmin = 0x7FFFFFFF, rmax = (int32_t)0x80000000;
#pragma unroll
    for (int k = 0; k < MAX_RADII_K; ++k) {
        rr[k] = k < (int)nradii ? radii[k] : 0x7FFFFFFF;
        if (k < (int)nradii) {
            rmin = min(rmin, rr[k]);
            rmax = max(rmax, rr[k]);
        }
    }
    const bool table = rmin >= 0 && rmax <= (int32_t)COUNT_LUT - 2;
    const int32_t *row = rowoff ? S + rowoff[q] : S + (size_t)q * ldS;
    const u32 n = rowlen ? rowlen[q] : nall;
    const bool vec = (reinterpret_cast<uintptr_t>(row) & 15) == 0;
    const u32 n4 = vec ? n / 4 : 0;
    u32 c[MAX_RADII_K];
#pragma unroll
    for (int k = 0; k < MAX_RADII_K; ++k) c[k] = 0;
    if (table) {
        // entry e holds score e - 1 (entry 0: every negative score, idx 0)
        for (u32 e = tid; e < (u32)rmax + 2; e += 256) {
            const int32_t x = (int32_t)e - 1;
            u32 idx = 0;
#pragma unroll
            for (int k = 0; k < MAX_RADII_K; ++k) idx += rr[k] <= x;
            u32 w[4] = {0, 0, 0, 0};
            if (idx) {
                const u32 f = idx - 1;
#pragma unroll
                for (int j = 0; j < 4; ++j) w[j] = (f >> 2) == (u32)j ? 1u << (8 * (f & 3)) : 0u;
            }
            lds4[e] = make_uint4(w[0], w[1], w[2], w[3]);
        }
        __syncthreads();
        u32 h0 = 0, h1 = 0, h2 = 0, h3 = 0, pend = 0;
        auto flush = [&] {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                c[b] += (h0 >> (8 * b)) & 0xFF;
                c[4 + b] += (h1 >> (8 * b)) & 0xFF;
                c[8 + b] += (h2 >> (8 * b)) & 0xFF;
                c[12 + b] += (h3 >> (8 * b)) & 0xFF;
            }
            h0 = h1 = h2 = h3 = 0;
            pend = 0;
        };
        auto add = [&](int32_t v) {
            const uint4 t = lds4[min(max(v, -1), rmax) + 1];
            h0 += t.x;
            h1 += t.y;
            h2 += t.z;
            h3 += t.w;
        };
        const int4 *r4 = reinterpret_cast<const int4 *>(row);
        for (u32 j0 = 0; j0 < n4; j0 += 1024) {
            int4 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const u32 j = j0 + 256 * u + tid;
                v[u] = j < n4 ? r4[j] : make_int4(-1, -1, -1, -1);  // -1: table entry 0, no field
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                add(v[u].x);
                add(v[u].y);
                add(v[u].z);
                add(v[u].w);
            }
            pend += 16;
            if (pend > 255 - 16) flush();
        }
        for (u32 j = 4 * n4 + tid; j < n; j += 256) {
            add(row[j]);
            if (++pend == 255) flush();
        }
        flush();
    } else {
        auto cmp = [&](int32_t v) {
#pragma unroll
            for (int k = 0; k < MAX_RADII_K; ++k) c[k] += v >= rr[k];
        };
        const int4 *r4 = reinterpret_cast<const int4 *>(row);
        for (u32 j = tid; j < n4; j += 256) {
            const int4 v = r4[j];
            cmp(v.x);
            cmp(v.y);
            cmp(v.z);
            cmp(v.w);
        }
        for (u32 j = 4 * n4 + tid; j < n; j += 256) cmp(row[j]);
    }
    // sum the 16 counters of the 256 threads: transpose through LDS (17-word
    // rows), 16 threads per counter add 16 rows each, then a 16-lane reduction
    __syncthreads();
    u32 *tr = reinterpret_cast<u32 *>(lds4);
#pragma unroll
    for (int k = 0; k < MAX_RADII_K; ++k) tr[tid * 17 + k] = c[k];
    __syncthreads();
    const u32 f = tid >> 4, part = tid & 15;
    u32 x = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) x += tr[(part * 16 + i) * 17 + f];
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if (part == 0) tot[f] = x;
    __syncthreads();
    if (tid < nradii) {
        u32 v;
        if (table) {
            const int32_t mine = radii[tid];
            u32 need = 0;  // idx of a score equal to this radius
#pragma unroll
            for (int k = 0; k < MAX_RADII_K; ++k) need += rr[k] <= mine;
            v = 0;
            for (u32 g = need - 1; g < MAX_RADII_K; ++g) v += tot[g];
        } else {
            v = tot[tid];
        }
        counts[(size_t)tid * R + q] = v;
    }
}

// One workgroup of 16 waves: wave w sums the R counts of radii w, w + 16, ...,
// then the first radius whose sum is at most smax is chosen.
// (the queries are the Rq columns from col0 of the R columns per radius)
__device__ __forceinline__ void choose_body(const u32 *counts, u32 R, u32 col0, u32 Rq, const int32_t *radii,
                                            u32 nradii, u64 smax, int any, int32_t *out)
{
    __shared__ u64 sums[MAX_ROUND_RADII];
    const u32 lane = threadIdx.x & 63;
    for (u32 k = threadIdx.x >> 6; k < nradii; k += 16) {
        u64 c = 0;
        const u32 *row = counts + (size_t)k * R + col0;
        for (u32 q0 = 0; q0 < Rq; q0 += 512) {  // eight loads in flight per lane
            u32 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const u32 q = q0 + 64 * u + lane;
                v[u] = q < Rq ? row[q] : 0u;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) c += v[u];
        }
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
        if (lane == 0) sums[k] = c;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    int32_t r = 0x7FFFFFFF, kk = -1;
    if (any)
        for (u32 j = 0; j < nradii; ++j)
            if (sums[j] <= smax) {
                r = radii[j];
                kk = (int32_t)j;
                break;
            }
    out[0] = r;
    out[1] = kk;
}

__global__ __launch_bounds__(1024) void k_autosplit_choose(const u32 *counts, u32 R, const int32_t *radii,
                                                           u32 nradii, u64 smax, int any, int32_t *out)
{
    choose_body(counts, R, 0, R, radii, nradii, smax, any, out);
}

// One workgroup per group.
__global__ __launch_bounds__(1024) void k_autosplit_choose_groups(const u32 *counts, u32 R, const u32 *qoff,
                                                                  const int32_t *radii, u32 nradii, const u64 *smax,
                                                                  const int32_t *any, int32_t *out)
{
    const u32 g = blockIdx.x;
    choose_body(counts, R, qoff[g], qoff[g + 1] - qoff[g], radii, nradii, smax[g], any[g], out + 2 * g);
}

// Copies between device memory and mapped pinned host memory inside a kernel,
// so a round trip needs no copy engine and no cross-engine wait.
__global__ void k_copy_words(const u32 *src, u32 *dst, u32 n)
{
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[i];
}

__global__ void k_publish(const int32_t *choice, const u64 *total, const u32 *counts, u32 ncounts, u32 *out)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out[0] = (u32)choice[0];
        out[1] = (u32)choice[1];
        const u64 t = *total;
        out[2] = (u32)t;
        out[3] = (u32)(t >> 32);
    }
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < ncounts; i += gridDim.x * blockDim.x)
        out[4 + i] = counts[i];
}

__global__ void k_row_size(const u32 *mat, u32 n, u32 ldw, u32 ntip, u32 *size)
{
    const u32 lane = threadIdx.x & 63;
    const u32 row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= n) return;
    const u32 *m = mat + (size_t)row * ldw;
    u32 c = 0;
    for (u32 w = lane; w < ldw; w += 64) c += (u32)__builtin_popcount(m[w]);
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane == 0) size[row] = c > ntip - c ? ntip - c : c;
}

template <class T>
__global__ void k_fill(T *p, u64 count, T value)
{
    for (u64 x = blockIdx.x * (u64)blockDim.x + threadIdx.x; x < count; x += (u64)gridDim.x * blockDim.x)
        p[x] = value;
}

u32 grid_for(u64 items, u32 block) { return (u32)std::min<u64>((items + block - 1) / block, 65535ull * 8); }

}  // namespace

void launch_transpose_wordmajor(const u32 *in, u32 W, u32 n, u32 *out, u32 ldw, hipStream_t s)
{
    dim3 g(ceil_div(n, 32), ceil_div(ldw, 32));
    hipLaunchKernelGGL(k_transpose_wordmajor, g, dim3(256), 0, s, in, W, n, out, ldw);
    CASPHY_HIP(hipGetLastError());
}

void launch_pad_rows(const u32 *in, u32 W, u32 n, u32 *out, u32 ldw, hipStream_t s)
{
    hipLaunchKernelGGL(k_pad_rows, dim3(grid_for((u64)n * ldw, 256)), dim3(256), 0, s, in, W, n, out, ldw);
    CASPHY_HIP(hipGetLastError());
}

void launch_dna_detect(const u32 *raw, u64 nwords, u32 *flag, hipStream_t s)
{
    hipLaunchKernelGGL(k_dna_detect, dim3(grid_for(nwords, 256)), dim3(256), 0, s, raw, nwords, flag);
    CASPHY_HIP(hipGetLastError());
}

void launch_dna_pack2(const u32 *raw, u32 n, u32 nwRaw, u32 *out, u32 ldw, hipStream_t s)
{
    hipLaunchKernelGGL(k_dna_pack2, dim3(grid_for((u64)n * (ldw / 3), 256)), dim3(256), 0, s, raw, n, nwRaw, out, ldw);
    CASPHY_HIP(hipGetLastError());
}

void launch_dna_pack_nib(const u32 *raw, u32 n, u32 nwRaw, u32 *out, u32 ldw, hipStream_t s)
{
    hipLaunchKernelGGL(k_dna_pack_nib, dim3(grid_for((u64)n * ldw, 256)), dim3(256), 0, s, raw, n, nwRaw, out, ldw);
    CASPHY_HIP(hipGetLastError());
}

void launch_row_popc(const u32 *mask, u32 n, u32 mstride, u32 words, u32 *counts, hipStream_t s,
                     const int32_t *skip)
{
    hipLaunchKernelGGL(k_row_popc, dim3(ceil_div(n, 4)), dim3(256), 0, s, mask, n, mstride, words, counts, skip,
                       RowMap{});
    CASPHY_HIP(hipGetLastError());
}

void launch_row_popc_map(const u32 *mask, u32 n, const RowMap &map, u32 *counts, hipStream_t s)
{
    if (!n) return;
    hipLaunchKernelGGL(k_row_popc, dim3(ceil_div(n, 4)), dim3(256), 0, s, mask, n, 0u, 0u, counts, nullptr, map);
    CASPHY_HIP(hipGetLastError());
}

void launch_scan_counts(const u32 *counts, u32 n, u64 *offsets, hipStream_t s)
{
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, counts, n, offsets);
    CASPHY_HIP(hipGetLastError());
}

void launch_compact(const u32 *mask, u32 n, u32 mstride, u32 words, const u64 *offsets, u32 *idx, hipStream_t s,
                    u64 cap)
{
    hipLaunchKernelGGL(k_compact, dim3(ceil_div(n, CWAVES)), dim3(64 * CWAVES), 0, s, mask, n, mstride, words, offsets,
                       idx, cap, RowMap{});
    CASPHY_HIP(hipGetLastError());
}

void launch_compact_map(const u32 *mask, u32 n, const RowMap &map, const u64 *offsets, u32 *idx, hipStream_t s,
                        u64 cap)
{
    if (!n) return;
    hipLaunchKernelGGL(k_compact, dim3(ceil_div(n, CWAVES)), dim3(64 * CWAVES), 0, s, mask, n, 0u, 0u, offsets, idx,
                       cap, map);
    CASPHY_HIP(hipGetLastError());
}

void launch_count_radii(const int32_t *S, u32 R, u32 n, u64 ldS, const int32_t *radii, u32 nradii, u32 *counts,
                        hipStream_t s)
{
    if (!R) return;
    for (u32 g = 0; g < nradii; g += MAX_RADII_K) {  // one pass per group of 16 radii
        hipLaunchKernelGGL(k_count_radii, dim3(R), dim3(256), 0, s, S, R, n, ldS, radii + g,
                           std::min<u32>(MAX_RADII_K, nradii - g), counts + (size_t)g * R, nullptr, nullptr);
        CASPHY_HIP(hipGetLastError());
    }
}

void launch_count_radii_rows(const int32_t *S, u32 R, const u64 *rowoff, const u32 *rowlen, const int32_t *radii,
                             u32 nradii, u32 *counts, hipStream_t s)
{
    if (!R) return;
    for (u32 g = 0; g < nradii; g += MAX_RADII_K) {
        hipLaunchKernelGGL(k_count_radii, dim3(R), dim3(256), 0, s, S, R, 0u, (u64)0, radii + g,
                           std::min<u32>(MAX_RADII_K, nradii - g), counts + (size_t)g * R, rowoff, rowlen);
        CASPHY_HIP(hipGetLastError());
    }
}

void launch_autosplit_choose(const u32 *counts, u32 R, const int32_t *radii, u32 nradii, u64 smax, int any,
                             int32_t *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_autosplit_choose, dim3(1), dim3(1024), 0, s, counts, R, radii, nradii, smax, any, out);
    CASPHY_HIP(hipGetLastError());
}

void launch_autosplit_choose_groups(const u32 *counts, u32 R, const u32 *qoff, u32 ngroups, const int32_t *radii,
                                    u32 nradii, const u64 *smax, const int32_t *any, int32_t *out, hipStream_t s)
{
    if (!ngroups) return;
    hipLaunchKernelGGL(k_autosplit_choose_groups, dim3(ngroups), dim3(1024), 0, s, counts, R, qoff, radii, nradii,
                       smax, any, out);
    CASPHY_HIP(hipGetLastError());
}

void launch_copy_words(const u32 *src, u32 *dst, u32 n, hipStream_t s)
{
    if (!n) return;
    hipLaunchKernelGGL(k_copy_words, dim3(std::min<u32>(ceil_div(n, 256), 64)), dim3(256), 0, s, src, dst, n);
    CASPHY_HIP(hipGetLastError());
}

void launch_publish(const int32_t *choice, const u64 *total, const u32 *counts, u32 ncounts, u32 *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_publish, dim3(std::max<u32>(1, std::min<u32>(ceil_div(ncounts, 256), 64))), dim3(256), 0, s,
                       choice, total, counts, ncounts, out);
    CASPHY_HIP(hipGetLastError());
}

void launch_row_size(const u32 *mat, u32 n, u32 ldw, u32 ntip, u32 *size, hipStream_t s)
{
    if (!n) return;
    hipLaunchKernelGGL(k_row_size, dim3(ceil_div(n, 4)), dim3(256), 0, s, mat, n, ldw, ntip, size);
    CASPHY_HIP(hipGetLastError());
}

void launch_fill_u32(u32 *p, u64 count, u32 value, hipStream_t s)
{
    if (!count) return;
    hipLaunchKernelGGL(k_fill<u32>, dim3(grid_for(count, 256)), dim3(256), 0, s, p, count, value);
    CASPHY_HIP(hipGetLastError());
}

void launch_fill_u64(u64 *p, u64 count, u64 value, hipStream_t s)
{
    if (!count) return;
    hipLaunchKernelGGL(k_fill<u64>, dim3(grid_for(count, 256)), dim3(256), 0, s, p, count, value);
    CASPHY_HIP(hipGetLastError());
}

}  // namespace casphy
