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

CladePlan plan_clades(const int32_t *anc, const int32_t *des, u32 nedge, u32 nnode, u32 ntip)
{
    CASPHY_REQUIRE(ntip >= 1 && nnode >= 1, CASPHY_EARG, "clade bitsets: need tips and internal nodes");
    const u64 nall = (u64)ntip + nnode;
    CladePlan p;
    p.nbin = ((ntip - 1) >> 5) + 1;
    std::vector<int32_t> size(nnode, 0);
    std::vector<unsigned char> entered(nnode, 0), child(nall + 1, 0);
    for (u32 e = 0; e < nedge; ++e) {
        const int64_t a = anc[e], d = des[e];
        CASPHY_REQUIRE(a > ntip && (u64)a <= nall && d >= 1 && (u64)d <= nall, CASPHY_EARG,
                       "clade bitsets: edge " + std::to_string(e) + " names a node outside the tree");
        CASPHY_REQUIRE(!child[d], CASPHY_EARG, "clade bitsets: node " + std::to_string(d) + " has two parent edges");
        child[d] = 1;
    }
    int64_t root = -1;
    for (u64 v = (u64)ntip + 1; v <= nall; ++v)
        if (!child[v]) {
            CASPHY_REQUIRE(root < 0, CASPHY_EARG, "clade bitsets: more than one root");
            root = (int64_t)v;
        }
    CASPHY_REQUIRE(root > 0, CASPHY_EARG, "clade bitsets: no root");
    entered[root - ntip - 1] = 1;
    // every edge leaves a node whose own edge came earlier (the order in which
    // nodeDesBig's backward scan completes each row before writing it)
    std::vector<u32> nkid(nnode + 1, 0);
    for (u32 e = 0; e < nedge; ++e) {
        const u32 a = (u32)anc[e] - ntip - 1;
        CASPHY_REQUIRE(entered[a], CASPHY_EARG,
                       "clade bitsets: edge " + std::to_string(e) + " leaves a node before the edge into it");
        if ((u32)des[e] > ntip) entered[des[e] - ntip - 1] = 1;
        ++nkid[a + 1];
    }
    // children in edge order, then a depth-first walk: the tips below every
    // internal node are one run of the walk's tip order
    for (u32 v = 0; v < nnode; ++v) nkid[v + 1] += nkid[v];
    std::vector<u32> kid(nedge), fill(nkid.begin(), nkid.end() - 1);
    for (u32 e = 0; e < nedge; ++e) kid[fill[(u32)anc[e] - ntip - 1]++] = (u32)des[e];
    std::vector<u32> start(nnode, 0);
    std::vector<std::pair<u32, u32>> stack;  // (internal index, next child slot)
    stack.reserve(64);
    stack.push_back({(u32)(root - ntip - 1), nkid[root - ntip - 1]});
    start[root - ntip - 1] = 0;
    p.order.reserve(ntip);
    while (!stack.empty()) {
        auto &top = stack.back();
        if (top.second == nkid[top.first + 1]) {
            stack.pop_back();
            continue;
        }
        const u32 d = kid[top.second++];
        if (d <= ntip) {
            p.order.push_back(d - 1);
        } else {
            start[d - ntip - 1] = (u32)p.order.size();
            stack.push_back({d - ntip - 1, nkid[d - ntip - 1]});
        }
    }
    // backward pass: sizes (nodeDesBig's counts) and the rows in write order
    for (u32 e = nedge; e-- > 0;) {
        const u32 a = (u32)anc[e] - ntip - 1;
        const int32_t d = des[e];
        if ((u32)d <= ntip) {
            size[a] += 1;
        } else {
            const u32 di = (u32)d - ntip - 1;
            size[a] += size[di];
            p.lo.push_back(start[di]);
            p.cnt.push_back((u32)size[di]);
        }
    }
    p.size.resize(nnode > 0 ? nnode - 1 : 0);
    for (u32 i = 1; i < nnode; ++i) p.size[nnode - 1 - i] = size[i];
    return p;
}

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
