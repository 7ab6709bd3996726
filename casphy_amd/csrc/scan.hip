// scan.hip -- scan form of the query x cell scores (tryRadius with few queries).
//
// The reference scores one query row against every cell row per launch
// (cu/tryRadius.cu:26-50, one launch per replicate).  Here one pass streams
// every cell row from HBM once and scores it against up to SCAN_Q query rows
// held in LDS, so a pass is bound by HBM bandwidth: per 16 bytes of a cell row
// a lane does SCAN_Q x 4 AND + popcount pairs (at SCAN_Q = 4, about 40% of
// the VALU rate at full HBM bandwidth).
//   * one wave per cell row at a time, 64 lanes x 16 bytes per load and up to
//     eight loads in flight per lane, rows read with non-temporal loads (each
//     byte is used once);
//   * the SCAN_Q per-lane partial sums are reduced with DPP row shifts and
//     row broadcasts (no LDS traffic) and read out of lane 63;
//   * lane (q, k = l & 15) counts the rows whose score reaches radius k; the
//     workgroup's counts go to per-counter slots (counter-major, so each
//     counter's partials are contiguous) and a second launch of one
//     workgroup per counter sums them (atomics from every workgroup on the
//     same 64 counters would serialise on two cache lines).
#include <algorithm>

#include "kernels.hpp"

namespace casphy {

namespace {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr u32 SCAN_WAVES = 4;  // waves per workgroup share one copy of the queries in LDS

__global__ __launch_bounds__(64 * SCAN_WAVES) void k_scan_bit(const u32 *mat, u32 n, u32 ldw, const u32 *qrows, u32 Q,
                                                  const int32_t *radii, u32 nradii, u32 *part, int32_t *S)
{
    __shared__ u32 red[SCAN_WAVES][64];
    extern __shared__ uint4 qs[];  // SCAN_Q x ldw/4, zero rows for q >= Q
    const u32 v4 = ldw / 4;
    for (u32 x = threadIdx.x; x < SCAN_Q * v4; x += blockDim.x) {
        const u32 q = x / v4;
        qs[x] = q < Q ? reinterpret_cast<const uint4 *>(qrows)[x] : make_uint4(0, 0, 0, 0);
    }
    __syncthreads();
    const u32 lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const u32 myq = (lane >> 4) & 3, myk = lane & 15;
    const int32_t rk = myk < nradii ? radii[myk] : 0x7FFFFFFF;
    u32 cnt = 0;
    // eight 16-byte loads in flight per lane; past the row end a lane loads
    // nothing, scores zeros and reads a valid query word.  The first chunk of
    // the next row is loaded before the current row's reduction.
    u32x4 x[8];
    auto load_chunk = [&](u32 r, u32 w0) {
        const u32x4 *row = reinterpret_cast<const u32x4 *>(mat + (size_t)r * ldw);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const u32 w = w0 + 64 * u + lane;
            x[u] = w < v4 ? __builtin_nontemporal_load(row + w) : u32x4{0, 0, 0, 0};
        }
    };
    const u32 stride = gridDim.x * SCAN_WAVES;
    u32 r = blockIdx.x * SCAN_WAVES + wv;
    if (r < n) load_chunk(r, 0);
    for (; r < n; r += stride) {
        u32 p0 = 0, p1 = 0, p2 = 0, p3 = 0;
        for (u32 w0 = 0; w0 < v4; w0 += 512) {
            if (w0) load_chunk(r, w0);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const u32 w = w0 + 64 * u + lane;
                const u32 wq = w < v4 ? w : 0;
                const uint4 a = qs[wq], b = qs[v4 + wq], c = qs[2 * v4 + wq], d = qs[3 * v4 + wq];
                const u32x4 y = x[u];
                p0 += __builtin_popcount(y.x & a.x) + __builtin_popcount(y.y & a.y) +
                      __builtin_popcount(y.z & a.z) + __builtin_popcount(y.w & a.w);
                p1 += __builtin_popcount(y.x & b.x) + __builtin_popcount(y.y & b.y) +
                      __builtin_popcount(y.z & b.z) + __builtin_popcount(y.w & b.w);
                p2 += __builtin_popcount(y.x & c.x) + __builtin_popcount(y.y & c.y) +
                      __builtin_popcount(y.z & c.z) + __builtin_popcount(y.w & c.w);
                p3 += __builtin_popcount(y.x & d.x) + __builtin_popcount(y.y & d.y) +
                      __builtin_popcount(y.z & d.z) + __builtin_popcount(y.w & d.w);
            }
        }
        if (r + stride < n) load_chunk(r + stride, 0);
        const u32 t0 = __builtin_amdgcn_readlane(wave_inclusive_scan(p0), 63);
        const u32 t1 = __builtin_amdgcn_readlane(wave_inclusive_scan(p1), 63);
        const u32 t2 = __builtin_amdgcn_readlane(wave_inclusive_scan(p2), 63);
        const u32 t3 = __builtin_amdgcn_readlane(wave_inclusive_scan(p3), 63);
        const u32 v = myq == 0 ? t0 : myq == 1 ? t1 : myq == 2 ? t2 : t3;
        if (myq < Q) {
            cnt += (int32_t)v >= rk;
            if (S && myk == 0) S[(size_t)myq * n + r] = (int32_t)v;
        }
    }
    // per-workgroup counts, counter-major: part[counter * gridDim.x + workgroup]
    red[wv][lane] = cnt;
    __syncthreads();
    if (wv == 0) {
        u32 c = 0;
#pragma unroll
        for (u32 k = 0; k < SCAN_WAVES; ++k) c += red[k][lane];
        part[(size_t)lane * gridDim.x + blockIdx.x] = c;
    }
}

// counts[k*R + q] = sum over workgroups of counter (q, k)'s partials: one
// workgroup per counter, its partials contiguous, every load of a thread
// independent of the others (a single workgroup looping over all 64 counters
// was bound by the latency of its dependent loads).
__global__ __launch_bounds__(256) void k_scan_sum(const u32 *part, u32 nblocks, u32 Q, u32 nradii, u32 *counts,
                                                  u32 R)
{
    __shared__ u32 red[4];
    const u32 slot = blockIdx.x, q = (slot >> 4) & 3, k = slot & 15;
    if (q >= Q || k >= nradii) return;  // uniform per workgroup
    const u32 *p = part + (size_t)slot * nblocks;
    u32 c = 0;
#pragma unroll 8
    for (u32 b = threadIdx.x; b < nblocks; b += 256) c += p[b];
    c = __builtin_amdgcn_readlane(wave_inclusive_scan(c), 63);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) counts[(size_t)k * R + q] = red[0] + red[1] + red[2] + red[3];
}

}  // namespace

bool scan_fits(u32 ldw) { return ldw >= 256 && (size_t)SCAN_Q * ldw * 4 <= 64 * 1024; }

u32 scan_grid(u32 n) { return std::min<u32>(ceil_div(n, SCAN_WAVES), 2048); }

void launch_scan_bit(const u32 *mat, u32 n, u32 ldw, const u32 *qrows, u32 Q, const int32_t *radii, u32 nradii,
                     u32 *counts, u32 R, int32_t *S, u32 *part, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    if (!n || !Q) return;
    CASPHY_REQUIRE(Q <= SCAN_Q && Q <= R && scan_fits(ldw) && nradii <= 16, CASPHY_EINTERNAL,
                   "scan form: bad shape");
    const u32 grid = scan_grid(n);
    if (ev0) CASPHY_HIP(hipEventRecord(ev0, s));
    hipLaunchKernelGGL(k_scan_bit, dim3(grid), dim3(64 * SCAN_WAVES), (size_t)SCAN_Q * ldw * 4, s, mat, n, ldw, qrows,
                       Q, radii, nradii, part, S);
    CASPHY_HIP(hipGetLastError());
    if (ev1) CASPHY_HIP(hipEventRecord(ev1, s));
    hipLaunchKernelGGL(k_scan_sum, dim3(64), dim3(256), 0, s, part, grid, Q, nradii, counts, R);
    CASPHY_HIP(hipGetLastError());
}

}  // namespace casphy
