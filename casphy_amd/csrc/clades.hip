// clades.hip -- clade bitsets of a rooted tree (the bipartition producer).
//
// nodeDesBig (src/functions.cpp:1277-1366) scans R's edge matrix backwards on
// one host thread, OR-ing each finished child row into its parent's row in a
// bounded cache and writing a node's row when its parent edge is reached.  When
// every node's edge comes before its children's edges (cladewise order, as
// ape's read.tree and reorder(, "cladewise") give) each written row is the
// node's whole clade.  Here the host walks the tree once (depth first, children
// in edge order) so that the tips below every node form one run of the walk's
// tip order, and one workgroup per row builds the row in LDS from its run and
// writes it out; rows keep nodeDesBig's order.
#include <algorithm>
#include <utility>
#include <vector>

#include "clades.hpp"
#include "context.hpp"
#include "kernels.hpp"

namespace casphy {

namespace {

// Row r (at rows + r * stride): the bits of tips order[lo[r] .. lo[r] +
// cnt[r]) in nbin words, built in an LDS row when LDS, else by atomics on the
// pre-zeroed global row.
template <bool LDS>
__global__ __launch_bounds__(256) void k_clade_rows(const u32 *order, const u32 *lo, const u32 *cnt, u32 nbin,
                                                    u32 stride, u32 *rows)
{
    extern __shared__ u32 row[];
    const u32 r = blockIdx.x, tid = threadIdx.x;
    u32 *dst = rows + (size_t)r * stride;
    if (LDS) {
        for (u32 w = tid; w < nbin; w += 256) row[w] = 0u;
        __syncthreads();
    }
    const u32 b = lo[r], e = b + cnt[r];
    for (u32 k = b + tid; k < e; k += 256) {
        const u32 t = order[k];
        if (LDS)
            atomicOr(&row[t >> 5], 1u << (t & 31));
        else
            atomicOr(&dst[t >> 5], 1u << (t & 31));
    }
    if (LDS) {
        __syncthreads();
        for (u32 w = tid; w < nbin; w += 256) dst[w] = row[w];
    }
}

constexpr u32 LDS_ROW_MAX = 16384;  // words of one LDS row (64 KB): up to 524,288 tips

}  // namespace

void build_clade_rows(const CladePlan &p, u32 *rows, u32 stride, hipStream_t s)
{
    const u32 nrow = (u32)p.lo.size();
    if (!nrow) return;
    CASPHY_REQUIRE(stride >= p.nbin, CASPHY_EINTERNAL, "clade rows: stride below the row width");
    DevBuf dord(std::max<size_t>(p.order.size(), 1) * 4), dlo((size_t)nrow * 4), dcnt((size_t)nrow * 4);
    if (!p.order.empty()) h2d(dord.p, p.order.data(), p.order.size() * 4, s);
    h2d(dlo.p, p.lo.data(), (size_t)nrow * 4, s);
    h2d(dcnt.p, p.cnt.data(), (size_t)nrow * 4, s);
    const bool lds = p.nbin <= LDS_ROW_MAX;
    if (!lds || stride > p.nbin) launch_fill_u32(rows, (u64)nrow * stride, 0u, s);  // atomics / padding words
    if (lds)
        hipLaunchKernelGGL(k_clade_rows<true>, dim3(nrow), dim3(256), (size_t)p.nbin * 4, s, dord.as<u32>(),
                           dlo.as<u32>(), dcnt.as<u32>(), p.nbin, stride, rows);
    else
        hipLaunchKernelGGL(k_clade_rows<false>, dim3(nrow), dim3(256), 0, s, dord.as<u32>(), dlo.as<u32>(),
                           dcnt.as<u32>(), p.nbin, stride, rows);
    CASPHY_HIP(hipGetLastError());
}

}  // namespace casphy
