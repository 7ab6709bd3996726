// nodekeep.hip -- nodeKeep: which internal nodes of a joined tree to keep.
//
// cu/nodeKeepBit.cu and cu/nodeKeepIndel.cu walk the edges of a tree in file
// order and, for each internal node they reach, upload its state row and count
// per outgroup row the node's mutations the outgroup lacks (one launch and one
// round trip per node); the node is kept when every outgroup lacks at least
// `cutoff` of them, and the subtree below a kept node is skipped.  Here one
// launch decides every internal node -- one workgroup per node, one lane per
// outgroup row -- and leaves one flag per node; the host then walks the edges
// with the reference's skip rule.
#include <vector>

#include "context.hpp"
#include "kernels.hpp"
#include "nodekeep.hpp"
#include "plan.hpp"

namespace casphy {

namespace {

// retain[v] = 1 when every outgroup row lacks at least `cutoff` of node v's
// mutations (bit rows: popcount(node & ~out); indel rows: sites with node
// state > 1 and an outgroup state >= 1 that differs).  One workgroup per node,
// one lane per outgroup: the outgroups are word-major (outT[w][j]) so a lane's
// reads are coalesced across the wave and the node's word is the same for
// every lane; blocks of 256 outgroups stop once one of them fails.
template <bool INDEL>
__global__ __launch_bounds__(256) void k_node_retain(const u32 *state, u32 width, const u32 *outT, u32 nout,
                                                     u32 cutoff, u32 *retain)
{
    __shared__ u32 fail;
    const u32 v = blockIdx.x, tid = threadIdx.x;
    if (tid == 0) fail = 0;
    __syncthreads();
    const u32 *row = state + (size_t)v * width;
    for (u32 j0 = 0; j0 < nout; j0 += 256) {
        const u32 j = j0 + tid;
        bool miss = false;
        if (j < nout) {
            u32 c = 0;
#pragma unroll 8
            for (u32 w = 0; w < width; ++w) {
                const u32 a = row[w], b = outT[(size_t)w * nout + j];
                if (INDEL)
                    c += a > 1u && b != a && b >= 1u;
                else
                    c += (u32)__builtin_popcount(a & ~b);
            }
            miss = c < cutoff;
        }
        if (__ballot(miss) && (tid & 63) == 0) atomicOr(&fail, 1u);
        __syncthreads();
        const bool stop = fail != 0;
        __syncthreads();  // every lane has read `fail` before the next block may write it
        if (stop) break;
    }
    if (tid == 0) retain[v] = fail ? 0u : 1u;
}

}  // namespace

// Kept internal node ids (file order of the edges) of nodeKeepBit (indel =
// false) or nodeKeepIndel.  edge: nedge x 2 (0-based ids, tips 0..nnode);
// state: nnode x width; outmat: nout x width.  Returns the count.
u32 node_keep(bool indel, const u32 *edge, u32 nedge, const u32 *state, u32 nnode, u32 width, const u32 *outmat,
              u32 nout, u32 cutoff, u32 *kept, hipStream_t s)
{
    const u32 ntip = nnode + 1;
    std::vector<u32> retain(nnode, 1u);
    if (nnode && nout && width) {
        DevBuf dstate((size_t)nnode * width * 4), dout((size_t)nout * width * 4), doutT((size_t)nout * width * 4),
            dret((size_t)nnode * 4);
        h2d(dstate.p, state, (size_t)nnode * width * 4, s);
        h2d(dout.p, outmat, (size_t)nout * width * 4, s);
        // [nout][width] -> [width][nout]
        launch_transpose_wordmajor(dout.as<u32>(), nout, width, doutT.as<u32>(), nout, s);
        if (indel)
            hipLaunchKernelGGL(k_node_retain<true>, dim3(nnode), dim3(256), 0, s, dstate.as<u32>(), width,
                               doutT.as<u32>(), nout, cutoff, dret.as<u32>());
        else
            hipLaunchKernelGGL(k_node_retain<false>, dim3(nnode), dim3(256), 0, s, dstate.as<u32>(), width,
                               doutT.as<u32>(), nout, cutoff, dret.as<u32>());
        CASPHY_HIP(hipGetLastError());
        d2h(retain.data(), dret.p, (size_t)nnode * 4, s);
    } else if (nnode && nout && cutoff > 0) {
        retain.assign(nnode, 0u);  // empty rows: no mutation to find
    }
    const u32 nk = node_keep_walk(edge, nedge, retain.data(), ntip, kept);
    return nk;
}

}  // namespace casphy
