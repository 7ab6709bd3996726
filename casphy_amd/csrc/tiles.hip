// tiles.hip -- exact integer pair-scoring tile kernel with fused epilogues.
//
// Work decomposition (gfx950, wave64):
//   * a workgroup of 256 threads (4 waves) owns a 128-row tile of A rows
//     against a tile of 16 CB B rows (CB = 8: 128 x 128; CB = 4: 128 x 64, the
//     callers' 128 x 128 tile run as two column halves); thread (ty, tx) of the
//     16 x 16 thread grid owns rows ty*8..ty*8+7 and columns tx*CB..tx*CB+CB-1,
//     i.e. 8 CB exact u32 accumulators;
//   * row words are staged through LDS k-major (As[k][row]) in stages of KC
//     words, double buffered, so each thread reads its 8 rows / CB columns of
//     a word with 16-byte LDS reads and spends 2 VALU ops (AND/XOR + v_bcnt
//     accumulate) per pair-word;
//   * (CB, KC) -- the tile shape and the words each lane stages per step -- is
//     picked per shape class by an autotuner that times every candidate on the
//     real data (every tuned epilogue is idempotent);
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

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

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

// A variant: CB columns per thread (8: 128 x 128 tiles, 4: 128 x 64) and KC
// words per stage.  STEP words make one pair-word (3 for the 2-bit DNA form).
template <int METRIC, int KC, int CB>
struct Shape {
    static constexpr int STEP = METRIC == M_DNA2 ? 3 : 1;
    static constexpr int TC = 16 * CB;        // B rows per tile
    static constexpr int WA = KC / 2;         // words per thread per A row (2 threads a row)
    static constexpr int WB = KC * TC / 256;  // words per thread per B row
    static constexpr int V4A = WA / 4, V4B = WB / 4;
    static_assert(KC % STEP == 0 && WA % 4 == 0 && WB % 4 == 0, "stage width must split into 16-byte loads");
};

template <int METRIC, int EPI, int KC, int CB>
__global__ __launch_bounds__(NTHR) void k_tiles(TileArgs a)
{
    using SH = Shape<METRIC, KC, CB>;
    constexpr int TC = SH::TC, NSUB = TILE / TC;
    __shared__ __attribute__((aligned(16))) u32 As[2][KC][TILE];
    __shared__ __attribute__((aligned(16))) u32 Bs[2][KC][TC];

    const u32 nsub = a.ntiles * NSUB;
    const u32 per = (nsub + 7) / 8;
    const u32 ts = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (ts >= nsub) return;
    // a device threshold of INT32_MAX (autoSplit round with no qualifying
    // radius): nothing to score; the row counts skip the mask as well
    if (EPI == E_MASK && a.thr_dev && *a.thr_dev == 0x7FFFFFFF) return;
    u32 I, J;
    tile_coords(a, ts / NSUB, I, J);
    const u32 col0 = J * TILE + (ts % NSUB) * TC;

    const u32 tid = threadIdx.x;
    const u32 ty = tid >> 4, tx = tid & 15;

    // ---- loader setup: a thread loads WA words of one A row and WB of one B row
    const u32 lrowA = tid & (TILE - 1), lpartA = tid / TILE;
    const u32 lrowB = tid & (TC - 1), lpartB = tid / TC;
    const u32 ga = I * TILE + lrowA, gb = col0 + lrowB;
    const bool va = ga < a.nA, vb = gb < a.nB;
    const u32 *arow = nullptr, *brow = nullptr;
    if (va) arow = a.A + (size_t)(a.Aidx ? a.Aidx[ga] : ga) * a.ldw;
    if (vb) brow = a.B + (size_t)gb * a.ldw;

    const u32 nplanes = a.nplanes ? a.nplanes : 1;
    const u32 stages_per_plane = a.kw / KC;
    const u32 nstages = stages_per_plane * nplanes;

    uint4 ra[SH::V4A], rb[SH::V4B];
    auto gload = [&](u32 s) {
        const u32 plane = s / stages_per_plane;
        const u32 kbase = (s - plane * stages_per_plane) * KC;
        const u32 ka = kbase + lpartA * SH::WA, kb = kbase + lpartB * SH::WB;
#pragma unroll
        for (int v = 0; v < SH::V4A; ++v)
            ra[v] = va ? *reinterpret_cast<const uint4 *>(arow + ka + 4 * v) : make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < SH::V4B; ++v)
            rb[v] = vb ? *reinterpret_cast<const uint4 *>(brow + kb + 4 * v) : make_uint4(0, 0, 0, 0);
        if (METRIC == M_AND && a.nplanes) {
            const u32 *pm = a.pmask + (size_t)plane * a.kw + ka;
#pragma unroll
            for (int v = 0; v < SH::V4A; ++v) {
                uint4 m = *reinterpret_cast<const uint4 *>(pm + 4 * v);
                ra[v].x &= m.x; ra[v].y &= m.y; ra[v].z &= m.z; ra[v].w &= m.w;
            }
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int v = 0; v < SH::V4A; ++v) {
            const int k = lpartA * SH::WA + 4 * v;
            As[buf][k + 0][lrowA] = ra[v].x; As[buf][k + 1][lrowA] = ra[v].y;
            As[buf][k + 2][lrowA] = ra[v].z; As[buf][k + 3][lrowA] = ra[v].w;
        }
#pragma unroll
        for (int v = 0; v < SH::V4B; ++v) {
            const int k = lpartB * SH::WB + 4 * v;
            Bs[buf][k + 0][lrowB] = rb[v].x; Bs[buf][k + 1][lrowB] = rb[v].y;
            Bs[buf][k + 2][lrowB] = rb[v].z; Bs[buf][k + 3][lrowB] = rb[v].w;
        }
    };

    u32 acc[RB][CB];
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int c = 0; c < CB; ++c) acc[r][c] = 0;

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
                for (int c = 0; c < CB; ++c) acc[r][c] <<= sh;
        }
#pragma unroll 4
        for (int k = 0; k < KC; k += SH::STEP) {
            u32 av[SH::STEP][RB], bv[SH::STEP][CB];
#pragma unroll
            for (int q = 0; q < SH::STEP; ++q) {
                const uint4 x0 = *reinterpret_cast<const uint4 *>(&As[buf][k + q][ty * RB]);
                const uint4 x1 = *reinterpret_cast<const uint4 *>(&As[buf][k + q][ty * RB + 4]);
                av[q][0] = x0.x; av[q][1] = x0.y; av[q][2] = x0.z; av[q][3] = x0.w;
                av[q][4] = x1.x; av[q][5] = x1.y; av[q][6] = x1.z; av[q][7] = x1.w;
#pragma unroll
                for (int c4 = 0; c4 < CB / 4; ++c4) {
                    const uint4 y = *reinterpret_cast<const uint4 *>(&Bs[buf][k + q][tx * CB + 4 * c4]);
                    bv[q][4 * c4 + 0] = y.x; bv[q][4 * c4 + 1] = y.y; bv[q][4 * c4 + 2] = y.z; bv[q][4 * c4 + 3] = y.w;
                }
            }
#pragma unroll
            for (int r = 0; r < RB; ++r)
#pragma unroll
                for (int c = 0; c < CB; ++c) {
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
            for (int c = 0; c < CB; ++c) acc[r][c] <<= a.ptail;
    }

    const u32 i0 = I * TILE + ty * RB, j0 = col0 + tx * CB;

    if (EPI == E_MASK) {
        const int32_t thr = a.thr_dev ? *a.thr_dev : a.thr;
        // direct words: row i0 + r, word j0 / 32 holds this thread's CB columns
        // at bit j0 % 32; the TPW threads of one word combine by shuffles and
        // one of them, chosen by r, stores it
        constexpr u32 TPW = 32 / CB;
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            u32 bits = 0;
#pragma unroll
            for (int c = 0; c < CB; ++c)
                if ((int32_t)acc[r][c] >= thr && j0 + c < a.nB) bits |= 1u << c;
            u32 piece = bits << ((tx % TPW) * CB);
#pragma unroll
            for (u32 o = 1; o < TPW; o <<= 1) piece |= __shfl_xor(piece, o);
            if ((tx % TPW) == (u32)r * TPW / RB && i0 + r < a.nA)
                a.mask[(size_t)(i0 + r) * a.mstride + (j0 >> 5)] = piece;
        }
        if (a.triangle && I != J) {
            // mirrored words: row j0 + c, word i0 / 32 from the four ty rows of a wave
#pragma unroll
            for (int c = 0; c < CB; ++c) {
                u32 byte = 0;
#pragma unroll
                for (int r = 0; r < RB; ++r)
                    if ((int32_t)acc[r][c] >= thr && i0 + r < a.nA) byte |= 1u << r;
                u32 piece = byte << ((ty & 3) * 8);
                piece |= __shfl_xor(piece, 16);
                piece |= __shfl_xor(piece, 32);
                if ((ty & 3) == (u32)c * 4 / CB && j0 + c < a.nB)
                    a.mask[(size_t)(j0 + c) * a.mstride + (i0 >> 5)] = piece;
            }
        }
    } else if (EPI == E_SCORES) {
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int c = 0; c < CB; ++c) {
                const u32 i = i0 + r, j = j0 + c;
                if (i < a.nA && j < a.nB) {
                    a.S[(size_t)i * a.ldS + j] = (int32_t)acc[r][c];
                    if (a.triangle && I != J) a.S[(size_t)j * a.ldS + i] = (int32_t)acc[r][c];
                }
            }
    } else if (EPI == E_HIST) {
        u32 *hist = &As[0][0][0];  // TILE x (MAX_RADII + 1)
        constexpr int HB = MAX_RADII + 1;
        static_assert(2 * KC * TILE >= TILE * HB, "histogram does not fit the stage buffer");
        for (u32 x = tid; x < TILE * HB; x += NTHR) hist[x] = 0;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            if (i0 + r >= a.nA) continue;
#pragma unroll
            for (int c = 0; c < CB; ++c) {
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
            for (int c = 0; c < CB; ++c) {
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
        u32 cmin[CB];
#pragma unroll
        for (int c = 0; c < CB; ++c) cmin[c] = 0xFFFFFFFFu;
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            u32 rmin = 0xFFFFFFFFu;
            const bool vi = i0 + r < a.nA;
#pragma unroll
            for (int c = 0; c < CB; ++c) {
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
        for (int c = 0; c < CB; ++c) {
            u32 v = cmin[c];
            u32 o = __shfl_xor(v, 16); v = o < v ? o : v;
            o = __shfl_xor(v, 32); v = o < v ? o : v;
            if ((ty & 3) == 0 && j0 + c < a.nB) atomicMin(&a.bmin[j0 + c], v);
        }
    }
}

// Variants: (KC, CB) per metric, named kKCxCB; index 0 is the default.
constexpr int kNumValu = 4;
struct ValuShape {
    int kc, cb;
    const char *name;
};
ValuShape valu_shape(int metric, int v)
{
    if (metric == M_DNA2) {
        static const ValuShape d[kNumValu] = {{24, 8, "k24x8"}, {48, 8, "k48x8"}, {0, 0, "-"}, {48, 4, "k48x4"}};
        return d[v];
    }
    static const ValuShape b[kNumValu] = {{32, 8, "k32x8"}, {16, 8, "k16x8"}, {32, 4, "k32x4"}, {16, 4, "k16x4"}};
    return b[v];
}
bool valu_supported(int metric, int v, const TileArgs &a)
{
    const ValuShape sh = valu_shape(metric, v);
    return sh.kc > 0 && a.kw % (u32)sh.kc == 0;
}

template <int METRIC, int KC, int CB>
void launch_shape(int epi, const TileArgs &a, hipStream_t s)
{
    const u32 nsub = a.ntiles * (TILE / (16 * CB));
    const u32 grid = 8 * ((nsub + 7) / 8);
    if constexpr (METRIC == M_XOR) {  // Hamming distances feed the transfer-distance epilogue only
        CASPHY_REQUIRE(epi == E_TBE, CASPHY_EINTERNAL, "XOR metric only feeds the TBE epilogue");
        hipLaunchKernelGGL((k_tiles<METRIC, E_TBE, KC, CB>), dim3(grid), dim3(NTHR), 0, s, a);
        return;
    } else {
        switch (epi) {
        case E_MASK: hipLaunchKernelGGL((k_tiles<METRIC, E_MASK, KC, CB>), dim3(grid), dim3(NTHR), 0, s, a); break;
        case E_SCORES: hipLaunchKernelGGL((k_tiles<METRIC, E_SCORES, KC, CB>), dim3(grid), dim3(NTHR), 0, s, a); break;
        case E_HIST: hipLaunchKernelGGL((k_tiles<METRIC, E_HIST, KC, CB>), dim3(grid), dim3(NTHR), 0, s, a); break;
        case E_LOCALMAX:
            hipLaunchKernelGGL((k_tiles<METRIC, E_LOCALMAX, KC, CB>), dim3(grid), dim3(NTHR), 0, s, a);
            break;
        default: throw Error(CASPHY_EINTERNAL, "bad epilogue for metric");
        }
    }
}

template <int METRIC>
void launch_metric(int v, int epi, const TileArgs &a, hipStream_t s)
{
    if constexpr (METRIC == M_DNA2) {
        switch (v) {
        case 0: launch_shape<METRIC, 24, 8>(epi, a, s); break;
        case 1: launch_shape<METRIC, 48, 8>(epi, a, s); break;
        case 3: launch_shape<METRIC, 48, 4>(epi, a, s); break;
        default: throw Error(CASPHY_EINTERNAL, "VALU tiles: variant not built for the DNA form");
        }
    } else {
        switch (v) {
        case 0: launch_shape<METRIC, 32, 8>(epi, a, s); break;
        case 1: launch_shape<METRIC, 16, 8>(epi, a, s); break;
        case 2: launch_shape<METRIC, 32, 4>(epi, a, s); break;
        case 3: launch_shape<METRIC, 16, 4>(epi, a, s); break;
        default: throw Error(CASPHY_EINTERNAL, "VALU tiles: unknown variant");
        }
    }
}

void launch_variant(int v, int metric, int epi, const TileArgs &a, hipStream_t s)
{
    switch (metric) {
    case M_AND: launch_metric<M_AND>(v, epi, a, s); break;
    case M_XOR: launch_metric<M_XOR>(v, epi, a, s); break;
    case M_DNA2: launch_metric<M_DNA2>(v, epi, a, s); break;
    case M_NIB: launch_metric<M_NIB>(v, epi, a, s); break;
    default: throw Error(CASPHY_EINTERNAL, "bad metric");
    }
    CASPHY_HIP(hipGetLastError());
}

// CASPHY_VALU_VARIANT: force a variant by name (k32x8, k16x8, k32x4, k16x4
// for the bit metrics; k24x8, k48x8, k48x4 for the 2-bit DNA form); a name
// the metric does not have leaves its launches to the autotuner.
int valu_variant_from_env(int metric)
{
    const char *e = std::getenv("CASPHY_VALU_VARIANT");
    if (!e) return -1;
    for (int v = 0; v < kNumValu; ++v)
        if (valu_shape(metric, v).kc > 0 && std::strcmp(e, valu_shape(metric, v).name) == 0) return v;
    return -1;
}

// Autotuner: the first launch of each shape class (metric, epilogue,
// triangle, log2 of both row counts, words per row, weight planes) runs every
// candidate twice on the real data -- the mask, score, local-maximum and
// transfer-distance epilogues are idempotent (stores, atomic max / min of
// identical values) -- times the second run with events and keeps the
// fastest.  tryRadius counts (E_HIST, atomic adds) and launches under stream
// capture take the default.
struct ValuKey {
    int metric, epi, tri;
    u32 la, lb, kw, planes;
    bool operator<(const ValuKey &o) const
    {
        return std::tie(metric, epi, tri, la, lb, kw, planes) <
               std::tie(o.metric, o.epi, o.tri, o.la, o.lb, o.kw, o.planes);
    }
};
struct ValuEntry {
    int variant;
    float ms[kNumValu];
};
std::mutex g_valu_mu;
std::map<ValuKey, ValuEntry> g_valu;

u32 log2_ceil_u(u32 x)
{
    u32 l = 0;
    while ((1u << l) < x && l < 31) ++l;
    return l;
}

}  // namespace

void launch_tiles(int metric, int epi, const TileArgs &a, hipStream_t s)
{
    if (a.ntiles == 0) return;
    CASPHY_REQUIRE(a.kw > 0 && a.kw % stage_words(metric) == 0 && a.kw <= a.ldw, CASPHY_EINTERNAL,
                   "tile row width not a multiple of the stage width");
    CASPHY_REQUIRE(metric != M_XOR || epi == E_TBE, CASPHY_EINTERNAL, "XOR metric only feeds the TBE epilogue");
    const int forced = valu_variant_from_env(metric);
    if (forced >= 0 && valu_supported(metric, forced, a)) return launch_variant(forced, metric, epi, a, s);
    const char *at = std::getenv("CASPHY_AUTOTUNE");
    if (epi == E_HIST || (at && at[0] == '0')) return launch_variant(0, metric, epi, a, s);
    const ValuKey key{metric, epi, a.triangle, log2_ceil_u(a.nA), log2_ceil_u(a.nB), a.kw, a.nplanes};
    {
        std::lock_guard<std::mutex> lk(g_valu_mu);
        auto it = g_valu.find(key);
        if (it != g_valu.end()) return launch_variant(it->second.variant, metric, epi, a, s);
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    CASPHY_HIP(hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone) return launch_variant(0, metric, epi, a, s);
    // a radius mask whose threshold lives in device memory may score nothing
    // (INT32_MAX): the candidates are timed with a host threshold, then the
    // real launch follows with the choice
    TileArgs t = a;
    const bool relaunch = epi == E_MASK && a.thr_dev;
    if (relaunch) {
        t.thr_dev = nullptr;
        t.thr = 1;
    }
    ValuEntry e{};
    for (float &m : e.ms) m = -1.0f;
    e.variant = 0;
    hipEvent_t ev[2];
    CASPHY_HIP(hipEventCreate(&ev[0]));
    CASPHY_HIP(hipEventCreate(&ev[1]));
    float best = 1e30f;
    try {
        for (int v = 0; v < kNumValu; ++v) {
            if (!valu_supported(metric, v, a)) continue;
            launch_variant(v, metric, epi, t, s);  // untimed: first-use costs
            CASPHY_HIP(hipEventRecord(ev[0], s));
            launch_variant(v, metric, epi, t, s);
            CASPHY_HIP(hipEventRecord(ev[1], s));
            CASPHY_HIP(hipEventSynchronize(ev[1]));
            float ms = 0;
            CASPHY_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
            e.ms[v] = ms;
            if (ms < best) {
                best = ms;
                e.variant = v;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        throw;
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (relaunch) launch_variant(e.variant, metric, epi, a, s);
    std::lock_guard<std::mutex> lk(g_valu_mu);
    g_valu.emplace(key, e);
}

std::string valu_tuning_json()
{
    std::lock_guard<std::mutex> lk(g_valu_mu);
    std::string out = "[";
    bool first = true;
    for (const auto &kv : g_valu) {
        const ValuKey &k = kv.first;
        const ValuEntry &e = kv.second;
        char buf[200];
        std::snprintf(buf, sizeof buf, "%s{\"form\": \"valu\", \"metric\": %d, \"epilogue\": %d, \"triangle\": %d, \"log2_rows_a\": %u, "
                      "\"log2_rows_b\": %u, \"words\": %u, \"planes\": %u, ",
                      first ? "" : ", ", k.metric, k.epi, k.tri, k.la, k.lb, k.kw, k.planes);
        out += buf;
        out += std::string("\"chosen\": \"") + valu_shape(k.metric, e.variant).name + "\", \"ms\": {";
        bool f2 = true;
        for (int v = 0; v < kNumValu; ++v) {
            if (e.ms[v] < 0) continue;
            std::snprintf(buf, sizeof buf, "%s\"%s\": %.6f", f2 ? "" : ", ", valu_shape(k.metric, v).name, e.ms[v]);
            out += buf;
            f2 = false;
        }
        out += "}}";
        first = false;
    }
    return out + "]";
}

}  // namespace casphy
