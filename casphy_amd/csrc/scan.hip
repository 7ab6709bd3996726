// scan.hip -- scan form of the query x cell scores (tryRadius with few queries).
//
// The reference scores one query row against every cell row per launch
// (cu/tryRadius.cu:26-50, one launch per replicate).  Here one pass streams
// every cell row from HBM once and scores it against up to SCAN_Q query rows
// held in LDS, so a pass is bound by HBM bandwidth: per 16 bytes of a cell row
// a lane does SCAN_Q x 4 AND + popcount pairs (at SCAN_Q = 4, about 40% of
// the VALU rate at full HBM bandwidth).
//   * one wave per cell row at a time, 64 lanes x 16 bytes per step, rows
//     read with non-temporal loads (each byte is used once);
//   * the SCAN_Q per-lane partial sums are reduced with a transposed
//     butterfly: after the exchanges over lane bits 5 and 4 every 16-lane
//     group holds one query's partial sums, which the remaining four steps
//     reduce, so lane l ends with the score of query (l >> 4) & 3;
//   * lane (q, k = l & 15) counts the rows whose score reaches radius k, and
//     each wave adds its counts to the global counters once at the end.
#include <algorithm>

#include "kernels.hpp"

namespace casphy {

namespace {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_scan_bit(const u32 *mat, u32 n, u32 ldw, const u32 *qrows, u32 Q,
                                                  const int32_t *radii, u32 nradii, u32 *counts, u32 R,
                                                  int32_t *S)
{
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
    for (u32 r = blockIdx.x * 4 + wv; r < n; r += gridDim.x * 4) {
        const u32x4 *row = reinterpret_cast<const u32x4 *>(mat + (size_t)r * ldw);
        u32 p0 = 0, p1 = 0, p2 = 0, p3 = 0;
        auto acc = [&](const u32x4 x, u32 w) {
            const uint4 a = qs[w], b = qs[v4 + w], c = qs[2 * v4 + w], d = qs[3 * v4 + w];
            p0 += __builtin_popcount(x.x & a.x) + __builtin_popcount(x.y & a.y) + __builtin_popcount(x.z & a.z) +
                  __builtin_popcount(x.w & a.w);
            p1 += __builtin_popcount(x.x & b.x) + __builtin_popcount(x.y & b.y) + __builtin_popcount(x.z & b.z) +
                  __builtin_popcount(x.w & b.w);
            p2 += __builtin_popcount(x.x & c.x) + __builtin_popcount(x.y & c.y) + __builtin_popcount(x.z & c.z) +
                  __builtin_popcount(x.w & c.w);
            p3 += __builtin_popcount(x.x & d.x) + __builtin_popcount(x.y & d.y) + __builtin_popcount(x.z & d.z) +
                  __builtin_popcount(x.w & d.w);
        };
        u32 w = lane;
        for (; w + 192 < v4; w += 256) {  // four 16-byte loads in flight per lane
            const u32x4 x0 = __builtin_nontemporal_load(row + w), x1 = __builtin_nontemporal_load(row + w + 64),
                        x2 = __builtin_nontemporal_load(row + w + 128), x3 = __builtin_nontemporal_load(row + w + 192);
            acc(x0, w);
            acc(x1, w + 64);
            acc(x2, w + 128);
            acc(x3, w + 192);
        }
        for (; w < v4; w += 64) acc(__builtin_nontemporal_load(row + w), w);
        // lane bit 5 picks queries {0,1} or {2,3}; lane bit 4 then one of the pair
        const bool h5 = lane & 32, h4 = lane & 16;
        u32 k0 = h5 ? p2 : p0, k1 = h5 ? p3 : p1;
        const u32 s0 = h5 ? p0 : p2, s1 = h5 ? p1 : p3;
        k0 += __shfl_xor(s0, 32);
        k1 += __shfl_xor(s1, 32);
        u32 v = h4 ? k1 : k0;
        v += __shfl_xor(h4 ? k0 : k1, 16);
        v += __shfl_xor(v, 8);
        v += __shfl_xor(v, 4);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 1);
        if (myq < Q) {
            cnt += (int32_t)v >= rk;
            if (S && myk == 0) S[(size_t)myq * n + r] = (int32_t)v;
        }
    }
    if (myq < Q && myk < nradii && cnt) atomicAdd(&counts[(size_t)myk * R + myq], cnt);
}

}  // namespace

bool scan_fits(u32 ldw) { return ldw >= 256 && (size_t)SCAN_Q * ldw * 4 <= 64 * 1024; }

void launch_scan_bit(const u32 *mat, u32 n, u32 ldw, const u32 *qrows, u32 Q, const int32_t *radii, u32 nradii,
                     u32 *counts, u32 R, int32_t *S, hipStream_t s)
{
    if (!n || !Q) return;
    CASPHY_REQUIRE(Q <= SCAN_Q && scan_fits(ldw) && nradii <= 16, CASPHY_EINTERNAL, "scan form: bad shape");
    const u32 grid = std::min<u32>(ceil_div(n, 4), 2048);
    hipLaunchKernelGGL(k_scan_bit, dim3(grid), dim3(256), (size_t)SCAN_Q * ldw * 4, s, mat, n, ldw, qrows, Q, radii,
                       nradii, counts, R, S);
    CASPHY_HIP(hipGetLastError());
}

}  // namespace casphy
