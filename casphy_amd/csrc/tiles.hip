// tiles.hip -- exact integer pair-scoring tile kernel with fused epilogues.
//
// Work decomposition (gfx950, wave64):
//   * a workgroup of 256 threads (4 waves) owns a 128 x 128 tile of pairs;
//     thread (ty, tx) of the 16 x 16 thread grid owns rows ty*8..ty*8+7 and
//     columns tx*8..tx*8+7, i.e. 64 exact u32 accumulators;
//   * row words are staged through LDS k-major (As[k][row]) in 32-word (128 B,
//     one cache line per row) stages, double buffered, so each thread reads
//     its 8 rows / 8 columns of a word with two 16-byte LDS reads each and
//     spends 2 VALU ops (AND/XOR + v_bcnt_u32_b32 accumulate) per pair-word;
//   * tiles are enumerated (upper triangle for self-scoring) and remapped so
//     that consecutive tiles, which share their A rows, land on the same XCD
//     (workgroups are dealt round-robin over the 8 XCDs).
// Epilogues (all integer, no host round trip):
//   E_MASK      radius bitmask words M[i][j/32] for S(i,j) >= r, written for
//               both (i,j) and (j,i) from one upper-triangle tile; the 32-bit
//               words are assembled with lane shuffles (no atomics);
//   E_SCORES    the score matrix itself (and its mirror);
//   E_HIST      tryRadius counts for up to 16 radii via an LDS histogram;
//   E_LOCALMAX  per row max over j > i of (S, smallest j) as a packed u64 key;
//   E_TBE       min(H, ntip - H) row and column minima (XOR metric).
#include "tiles.hpp"

namespace casphy {

namespace {

constexpr int NTHR = 256;
constexpr int RB = 8;

__device__ __forceinline__ u32 pop(u32 x) { return (u32)__builtin_popcount(x); }

// Count of non-zero nibbles of x.
__device__ __forceinline__ u32 nib_nonzero(u32 x)
{
    x |= x >> 1;
    x |= x >> 2;
    return pop(x & 0x11111111u);
}

template <int METRIC>
struct Stage {
    static constexpr int KC = METRIC == M_DNA2 ? KC_DNA2 : KC_BIT;
    static constexpr int STEP = METRIC == M_DNA2 ? 3 : 1;
    static constexpr int HALF = KC / 2;      // words loaded per thread per operand
    static constexpr int V4 = HALF / 4;      // 16-byte loads per thread per operand
};

__device__ __forceinline__ void tile_coords(const TileArgs &a, u32 t, u32 &I, u32 &J)
{
    if (a.triangle) {
        u32 j = (u32)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
        while ((u64)j * (j + 1) / 2 > t) --j;
        while ((u64)(j + 1) * (j + 2) / 2 <= t) ++j;
        J = j;
        I = t - (u32)((u64)j * (j + 1) / 2);
    } else {
        I = t / a.ntB;
        J = t - I * a.ntB;
    }
}

template <int METRIC, int EPI>
__global__ __launch_bounds__(NTHR) void k_tiles(TileArgs a)
{
    using ST = Stage<METRIC>;
    constexpr int KC = ST::KC;
    __shared__ __attribute__((aligned(16))) u32 As[2][KC][TILE];
    __shared__ __attribute__((aligned(16))) u32 Bs[2][KC][TILE];

    const u32 per = (a.ntiles + 7) / 8;
    const u32 t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (t >= a.ntiles) return;
    // a device threshold of INT32_MAX (autoSplit round with no qualifying
    // radius): nothing to score; the row counts skip the mask as well
    if (EPI == E_MASK && a.thr_dev && *a.thr_dev == 0x7FFFFFFF) return;
    u32 I, J;
    tile_coords(a, t, I, J);

    const u32 tid = threadIdx.x;
    const u32 ty = tid >> 4, tx = tid & 15;

    // ---- loader setup: thread loads HALF words of one A row and one B row
    const u32 lrow = tid & (TILE - 1), lpart = tid >> 7;
    const u32 ga = I * TILE + lrow, gb = J * TILE + lrow;
    const bool va = ga < a.nA, vb = gb < a.nB;
    const u32 *arow = nullptr, *brow = nullptr;
    if (va) arow = a.A + (size_t)(a.Aidx ? a.Aidx[ga] : ga) * a.ldw;
    if (vb) brow = a.B + (size_t)gb * a.ldw;

    const u32 nplanes = a.nplanes ? a.nplanes : 1;
    const u32 stages_per_plane = a.kw / KC;
    const u32 nstages = stages_per_plane * nplanes;

    uint4 ra[ST::V4], rb[ST::V4];
    auto gload = [&](u32 s) {
        const u32 plane = s / stages_per_plane;
        const u32 k0 = (s - plane * stages_per_plane) * KC + lpart * ST::HALF;
#pragma unroll
        for (int v = 0; v < ST::V4; ++v) {
            ra[v] = va ? *reinterpret_cast<const uint4 *>(arow + k0 + 4 * v) : make_uint4(0, 0, 0, 0);
            rb[v] = vb ? *reinterpret_cast<const uint4 *>(brow + k0 + 4 * v) : make_uint4(0, 0, 0, 0);
        }
        if (METRIC == M_AND && a.nplanes) {
            const u32 *pm = a.pmask + (size_t)plane * a.kw + k0;
#pragma unroll
            for (int v = 0; v < ST::V4; ++v) {
                uint4 m = *reinterpret_cast<const uint4 *>(pm + 4 * v);
                ra[v].x &= m.x; ra[v].y &= m.y; ra[v].z &= m.z; ra[v].w &= m.w;
            }
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int v = 0; v < ST::V4; ++v) {
            const int k = lpart * ST::HALF + 4 * v;
            As[buf][k + 0][lrow] = ra[v].x; As[buf][k + 1][lrow] = ra[v].y;
            As[buf][k + 2][lrow] = ra[v].z; As[buf][k + 3][lrow] = ra[v].w;
            Bs[buf][k + 0][lrow] = rb[v].x; Bs[buf][k + 1][lrow] = rb[v].y;
            Bs[buf][k + 2][lrow] = rb[v].z; Bs[buf][k + 3][lrow] = rb[v].w;
        }
    };

    u32 acc[RB][RB];
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int c = 0; c < RB; ++c) acc[r][c] = 0;

    gload(0);
    lstore(0);
    __syncthreads();
    for (u32 s = 0; s < nstages; ++s) {
        const int buf = s & 1;
        if (s + 1 < nstages) gload(s + 1);
        if (a.nplanes && s > 0 && (s % stages_per_plane) == 0) {
            const u32 sh = a.pshift[s / stages_per_plane];
#pragma unroll
            for (int r = 0; r < RB; ++r)
#pragma unroll
                for (int c = 0; c < RB; ++c) acc[r][c] <<= sh;
        }
#pragma unroll 4
        for (int k = 0; k < KC; k += ST::STEP) {
            u32 av[ST::STEP][RB], bv[ST::STEP][RB];
#pragma unroll
            for (int q = 0; q < ST::STEP; ++q) {
                uint4 x0 = *reinterpret_cast<const uint4 *>(&As[buf][k + q][ty * RB]);
                uint4 x1 = *reinterpret_cast<const uint4 *>(&As[buf][k + q][ty * RB + 4]);
                uint4 y0 = *reinterpret_cast<const uint4 *>(&Bs[buf][k + q][tx * RB]);
                uint4 y1 = *reinterpret_cast<const uint4 *>(&Bs[buf][k + q][tx * RB + 4]);
                av[q][0] = x0.x; av[q][1] = x0.y; av[q][2] = x0.z; av[q][3] = x0.w;
                av[q][4] = x1.x; av[q][5] = x1.y; av[q][6] = x1.z; av[q][7] = x1.w;
                bv[q][0] = y0.x; bv[q][1] = y0.y; bv[q][2] = y0.z; bv[q][3] = y0.w;
                bv[q][4] = y1.x; bv[q][5] = y1.y; bv[q][6] = y1.z; bv[q][7] = y1.w;
            }
#pragma unroll
            for (int r = 0; r < RB; ++r)
#pragma unroll
                for (int c = 0; c < RB; ++c) {
                    if (METRIC == M_AND) {
                        acc[r][c] += pop(av[0][r] & bv[0][c]);
                    } else if (METRIC == M_XOR) {
                        acc[r][c] += pop(av[0][r] ^ bv[0][c]);
                    } else if (METRIC == M_NIB) {
                        acc[r][c] += nib_nonzero(av[0][r] & bv[0][c]);
                    } else {  // DNA2: (hi, lo, mask)
                        const u32 d = (av[0][r] ^ bv[0][c]) | (av[1][r] ^ bv[1][c]);
                        acc[r][c] += pop(av[2][r] & bv[2][c] & ~d);
                    }
                }
        }
        if (s + 1 < nstages) lstore(buf ^ 1);
        __syncthreads();
    }
    if (a.nplanes && a.ptail) {
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int c = 0; c < RB; ++c) acc[r][c] <<= a.ptail;
    }

    const u32 i0 = I * TILE + ty * RB, j0 = J * TILE + tx * RB;

    if (EPI == E_MASK) {
        const int32_t thr = a.thr_dev ? *a.thr_dev : a.thr;
        // direct words: row i, columns J*128 + 32w .. +31
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            u32 byte = 0;
#pragma unroll
            for (int c = 0; c < RB; ++c)
                if ((int32_t)acc[r][c] >= thr && j0 + c < a.nB) byte |= 1u << c;
            u32 piece = byte << ((tx & 3) * 8);
            piece |= __shfl_xor(piece, 1);
            piece |= __shfl_xor(piece, 2);
            if ((tx & 3) == (u32)(r >> 1) && i0 + r < a.nA)
                a.mask[(size_t)(i0 + r) * a.mstride + J * 4 + (tx >> 2)] = piece;
        }
        if (a.triangle && I != J) {
#pragma unroll
            for (int c = 0; c < RB; ++c) {
                u32 byte = 0;
#pragma unroll
                for (int r = 0; r < RB; ++r)
                    if ((int32_t)acc[r][c] >= thr && i0 + r < a.nA) byte |= 1u << r;
                u32 piece = byte << ((ty & 3) * 8);
                piece |= __shfl_xor(piece, 16);
                piece |= __shfl_xor(piece, 32);
                if ((ty & 3) == (u32)(c >> 1) && j0 + c < a.nB)
                    a.mask[(size_t)(j0 + c) * a.mstride + I * 4 + (ty >> 2)] = piece;
            }
        }
    } else if (EPI == E_SCORES) {
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int c = 0; c < RB; ++c) {
                const u32 i = i0 + r, j = j0 + c;
                if (i < a.nA && j < a.nB) {
                    a.S[(size_t)i * a.ldS + j] = (int32_t)acc[r][c];
                    if (a.triangle && I != J) a.S[(size_t)j * a.ldS + i] = (int32_t)acc[r][c];
                }
            }
    } else if (EPI == E_HIST) {
        u32 *hist = &As[0][0][0];  // TILE x (MAX_RADII + 1)
        constexpr int HB = MAX_RADII + 1;
        for (u32 x = tid; x < TILE * HB; x += NTHR) hist[x] = 0;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            if (i0 + r >= a.nA) continue;
#pragma unroll
            for (int c = 0; c < RB; ++c) {
                if (j0 + c >= a.nB) continue;
                u32 bin = 0;
                for (u32 k = 0; k < a.nradii; ++k) bin += (int32_t)acc[r][c] >= a.radii[k];
                if (bin) atomicAdd(&hist[(ty * RB + r) * HB + bin], 1u);
            }
        }
        __syncthreads();
        if (tid < TILE) {
            const u32 q = I * TILE + tid;
            if (q < a.nA) {
                u32 run = 0;
                for (int k = (int)a.nradii - 1; k >= 0; --k) {
                    run += hist[tid * HB + k + 1];
                    if (run) atomicAdd(&a.counts[(size_t)k * a.nA + q], run);
                }
            }
        }
    } else if (EPI == E_LOCALMAX) {
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const u32 i = i0 + r;
            u64 best = 0;
#pragma unroll
            for (int c = 0; c < RB; ++c) {
                const u32 j = j0 + c;
                if (j < a.nB && j > i) {
                    const u64 key = ((u64)acc[r][c] << 32) | (u64)(0xFFFFFFFFu - j);
                    best = key > best ? key : best;
                }
            }
#pragma unroll
            for (int m = 1; m < 16; m <<= 1) {
                const u64 o = __shfl_xor(best, m);
                best = o > best ? o : best;
            }
            if (tx == 0 && i < a.nA && best) atomicMax((unsigned long long *)&a.keys[i], (unsigned long long)best);
        }
    } else if (EPI == E_TBE) {
        u32 cmin[RB];
#pragma unroll
        for (int c = 0; c < RB; ++c) cmin[c] = 0xFFFFFFFFu;
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            u32 rmin = 0xFFFFFFFFu;
            const bool vi = i0 + r < a.nA;
#pragma unroll
            for (int c = 0; c < RB; ++c) {
                const u32 h = acc[r][c], hc = a.ntip - h;
                const u32 d = h < hc ? h : hc;
                if (vi && j0 + c < a.nB) {
                    rmin = d < rmin ? d : rmin;
                    cmin[c] = d < cmin[c] ? d : cmin[c];
                }
            }
#pragma unroll
            for (int m = 1; m < 16; m <<= 1) {
                const u32 o = __shfl_xor(rmin, m);
                rmin = o < rmin ? o : rmin;
            }
            if (tx == 0 && vi) atomicMin(&a.amin[i0 + r], rmin);
        }
#pragma unroll
        for (int c = 0; c < RB; ++c) {
            u32 v = cmin[c];
            u32 o = __shfl_xor(v, 16); v = o < v ? o : v;
            o = __shfl_xor(v, 32); v = o < v ? o : v;
            if ((ty & 3) == 0 && j0 + c < a.nB) atomicMin(&a.bmin[j0 + c], v);
        }
    }
}

template <int METRIC>
void launch_m(int epi, const TileArgs &a, hipStream_t s)
{
    const u32 grid = 8 * ((a.ntiles + 7) / 8);
    switch (epi) {
    case E_MASK: hipLaunchKernelGGL((k_tiles<METRIC, E_MASK>), dim3(grid), dim3(NTHR), 0, s, a); break;
    case E_SCORES: hipLaunchKernelGGL((k_tiles<METRIC, E_SCORES>), dim3(grid), dim3(NTHR), 0, s, a); break;
    case E_HIST: hipLaunchKernelGGL((k_tiles<METRIC, E_HIST>), dim3(grid), dim3(NTHR), 0, s, a); break;
    case E_LOCALMAX: hipLaunchKernelGGL((k_tiles<METRIC, E_LOCALMAX>), dim3(grid), dim3(NTHR), 0, s, a); break;
    default: throw Error(CASPHY_EINTERNAL, "bad epilogue for metric");
    }
}

}  // namespace

void launch_tiles(int metric, int epi, const TileArgs &a, hipStream_t s)
{
    if (a.ntiles == 0) return;
    CASPHY_REQUIRE(a.kw > 0 && a.kw % stage_words(metric) == 0 && a.kw <= a.ldw, CASPHY_EINTERNAL,
                   "tile row width not a multiple of the stage width");
    switch (metric) {
    case M_AND: launch_m<M_AND>(epi, a, s); break;
    case M_XOR:
        CASPHY_REQUIRE(epi == E_TBE, CASPHY_EINTERNAL, "XOR metric only feeds the TBE epilogue");
        hipLaunchKernelGGL((k_tiles<M_XOR, E_TBE>), dim3(8 * ((a.ntiles + 7) / 8)), dim3(NTHR), 0, s, a);
        break;
    case M_DNA2: launch_m<M_DNA2>(epi, a, s); break;
    case M_NIB: launch_m<M_NIB>(epi, a, s); break;
    default: throw Error(CASPHY_EINTERNAL, "bad metric");
    }
    CASPHY_HIP(hipGetLastError());
}

}  // namespace casphy
