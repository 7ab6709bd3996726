// tbe.hip -- transfer distances between the clades of whole trees by a
// dynamic program over one tree per lane group (no bitsets).
//
// For a reference clade y and a replicate clade x over the same ntip tips,
// TransferDisStream (cu/TransferDisStream.cu:22-105) takes the Hamming distance
// H of their bitsets, d = min(H, ntip - H), and keeps ymin[y] = min over x and
// xmin[x] = min over y.  With sizes sy, sx and I = |C_x & C_y|,
// H = sy + sx - 2 I, and for a fixed y the intersections I(x) of every node x
// of the replicate follow from one post-order pass over the replicate tree:
// a tip t contributes [t in C_y], an internal node the sum of its children.
// The tips below a reference node are one run [lo_y, lo_y + sy) of the
// reference tree's depth-first tip order, so [t in C_y] is one unsigned
// compare of t's position in that order.  One lane owns one clade (64 per
// wave), all lanes of a launch walk the same tree in the same order, and the
// walk is a program of uniform op words read through the scalar cache:
//   TIP p  push the running value, start a new one at [p in run]
//   ACC p  add [p in run]                       (a light tip child)
//   ADD    add the value under the top, pop     (a light internal child done)
//   ROW s  the running value is a clade row's   (its node, s tips, is done)
// and ADD may carry the ROW of the node it completes (flag + size).  At a
// row, d is evaluated against the lane's clade and the lane's minimum
// updated.  Children are walked heaviest subtree first, so a value is pushed
// only when the walk enters a light child (at most half its parent's tips):
// the per-lane stack, in LDS, holds at most log2(ntip) + 1 values.  ymin
// comes from walks of each replicate with the reference clades as lanes,
// xmin from walks of the reference with each replicate's clades as lanes --
// no atomics, no cross-lane reductions.  Per (clade, walk node) pair a
// handful of integer lane-ops: for 65,534 x 65,534 clades about a millisecond
// of VALU time per pass, against 5.6e14 ops of bitset products per replicate
// on the matrix cores.
#include <algorithm>
#include <thread>
#include <utility>
#include <vector>

#include "plan.hpp"
#include "context.hpp"
#include "kernels.hpp"
#include "tbe.hpp"

namespace casphy {

namespace {


struct WalkJob {
    u64 op_off;    // program: ops[op_off .. op_off + nops), nops a multiple of OP_BLOCK
    u64 lane_off;  // the lane tree's rows: lane_lo / lane_sz[lane_off .. + nlanes)
    u64 out_off;   // their minima: out[out_off .. + nlanes)
    u32 nops;
    u32 nlanes;
};

// One wave = 64 lanes x RPL clades per lane (rows first + r * 64 + lane);
// blockIdx.y = job.  Several clades per lane share each op's scalar decode
// (the uniform op stream costs scalar and branch issue slots that one clade's
// few vector ops cannot fill).  The stack is `depth` x RPL values per lane in
// dynamic LDS (WAVES x depth x RPL x 64 u32).
template <int RPL, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_tbe_walk(const u32 *ops, const u32 *lane_lo, const u32 *lane_sz,
                                                             const WalkJob *jobs, u32 ntip, u32 init, u32 depth,
                                                             u32 *out)
{
    extern __shared__ u32 stack[];
    const WalkJob jb = jobs[blockIdx.y];
    const u32 lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const u32 first = (blockIdx.x * WAVES + wv) * 64 * RPL;
    if (first >= jb.nlanes) return;  // wave-uniform
    u32 lo[RPL], sz[RPL], best[RPL];
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
        const u32 me = first + r * 64 + lane;
        const bool valid = me < jb.nlanes;
        lo[r] = valid ? lane_lo[jb.lane_off + me] : 0u;
        sz[r] = valid ? lane_sz[jb.lane_off + me] : 0u;  // 0: no tip is in the run
        best[r] = init;
    }
    const u32 *op = ops + jb.op_off;
    u32 *stk = stack + (size_t)wv * depth * RPL * 64 + lane;
    // the running counts are kept negated (neg = -I mod 2^32), so a row's
    // H = sz + s - 2 I is one three-operand add
    u32 neg[RPL];
#pragma unroll
    for (int r = 0; r < RPL; ++r) neg[r] = 0;
    u32 sp = 0;  // stack depth, uniform
    for (u32 i0 = 0; i0 < jb.nops; i0 += OP_BLOCK) {
        u32 wb[OP_BLOCK];
#pragma unroll
        for (u32 k = 0; k < OP_BLOCK; ++k) wb[k] = op[i0 + k];
#pragma unroll
        for (u32 k = 0; k < OP_BLOCK; ++k) {
            const u32 w = wb[k], arg = w & ARG_MASK;
            if (w & OP_POS) {
                if (w & OP_STACK) {  // TIP
#pragma unroll
                    for (int r = 0; r < RPL; ++r) {
                        stk[(sp * RPL + r) * 64] = neg[r];
                        neg[r] = (arg - lo[r]) < sz[r] ? 0xFFFFFFFFu : 0u;
                    }
                    sp = __builtin_amdgcn_readfirstlane(sp + 1);
                } else {  // ACC
#pragma unroll
                    for (int r = 0; r < RPL; ++r) neg[r] -= (arg - lo[r]) < sz[r] ? 1u : 0u;
                }
            } else {
                if (w & OP_STACK) {  // ADD
                    sp = __builtin_amdgcn_readfirstlane(sp - 1);
#pragma unroll
                    for (int r = 0; r < RPL; ++r) neg[r] += stk[(sp * RPL + r) * 64];
                }
                if (w & OP_ROW) {
#pragma unroll
                    for (int r = 0; r < RPL; ++r) {
                        const u32 h = sz[r] + arg + 2u * neg[r];
                        const u32 hc = ntip - h;
                        best[r] = min(best[r], min(h, hc));
                    }
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
        const u32 me = first + r * 64 + lane;
        if (me < jb.nlanes) out[jb.out_off + me] = best[r];
    }
}

template <int RPL>
void launch_walk_w(u32 waves, dim3 grid, size_t lds, hipStream_t s, const u32 *ops, const u32 *lo, const u32 *sz,
                   const WalkJob *jobs, u32 ntip, u32 init, u32 depth, u32 *out)
{
    if (waves == 4)
        hipLaunchKernelGGL((k_tbe_walk<RPL, 4>), grid, dim3(256), lds, s, ops, lo, sz, jobs, ntip, init, depth, out);
    else
        hipLaunchKernelGGL((k_tbe_walk<RPL, 1>), grid, dim3(64), lds, s, ops, lo, sz, jobs, ntip, init, depth, out);
}

void launch_walk(u32 rpl, u32 waves, dim3 grid, size_t lds, hipStream_t s, const u32 *ops, const u32 *lo,
                 const u32 *sz, const WalkJob *jobs, u32 ntip, u32 init, u32 depth, u32 *out)
{
    switch (rpl) {
    case 1: launch_walk_w<1>(waves, grid, lds, s, ops, lo, sz, jobs, ntip, init, depth, out); break;
    case 2: launch_walk_w<2>(waves, grid, lds, s, ops, lo, sz, jobs, ntip, init, depth, out); break;
    case 8: launch_walk_w<8>(waves, grid, lds, s, ops, lo, sz, jobs, ntip, init, depth, out); break;
    default: launch_walk_w<4>(waves, grid, lds, s, ops, lo, sz, jobs, ntip, init, depth, out); break;
    }
}

u32 clade_size(u32 cnt, u32 ntip) { return std::min(cnt, ntip - cnt); }

}  // namespace

void transfer_trees_dp(const int32_t *y_anc, const int32_t *y_des, u32 y_nedge, u32 y_nnode, const int32_t *x_anc,
                       const int32_t *x_des, const u32 *x_edge_offsets, const u32 *x_nnode, u32 nrep, u32 ntip,
                       u32 *ymin, u32 *ysize, u32 *xmin, u32 *xsize, hipStream_t s)
{
    CASPHY_REQUIRE(ntip >= 1 && ntip < NOT_IN_TREE, CASPHY_EARG, "transfer trees: 1 .. 2^29 - 1 tips");
    const TreeData ty = tree_data(y_anc, y_des, y_nedge, y_nnode, ntip);
    const u32 ny = (u32)ty.plan.lo.size();
    CASPHY_REQUIRE(ny > 0, CASPHY_EARG, "transfer trees: the reference tree has no clade rows");
    const u32 init = 32u * ty.plan.nbin;
    for (u32 y = 0; y < ny; ++y) ysize[y] = clade_size(ty.plan.cnt[y], ntip);

    // replicates in chunks: host planning of chunk k + 1 (threads over its
    // replicates) overlaps the device walks of chunk k
    const char *ce = std::getenv("CASPHY_TBE_CHUNK");
    const u32 chunk = std::max<u32>(1, std::min<u32>(ce ? (u32)std::atoi(ce) : 64u, 32767u));
    const u32 nthreads = std::max<u32>(1, std::min<u32>(16, std::thread::hardware_concurrency()));

    struct Chunk {
        u32 nb = 0;
        std::vector<u32> ops, lo, sz;
        std::vector<WalkJob> jobs;
        std::vector<u32> xsz;  // clade sizes of the replicates' rows, in order
        u64 nx = 0;            // rows of the chunk's replicates
        u32 depth = 1;         // deepest walk stack
        std::string err;
    };
    auto plan_chunk = [&](u32 b0, u32 nb, Chunk &c) {
        c = Chunk{};
        c.nb = nb;
        std::vector<TreeData> tx(nb);
        std::vector<std::string> errs(nthreads);
        std::vector<std::thread> pool;
        for (u32 th = 0; th < std::min(nthreads, nb); ++th)
            pool.emplace_back([&, th] {
                try {
                    for (u32 k = th; k < nb; k += nthreads) {
                        const u32 b = b0 + k;
                        CASPHY_REQUIRE(x_edge_offsets[b + 1] >= x_edge_offsets[b], CASPHY_EARG,
                                       "transfer trees: edge offsets must ascend");
                        tx[k] = tree_data(x_anc + x_edge_offsets[b], x_des + x_edge_offsets[b],
                                          x_edge_offsets[b + 1] - x_edge_offsets[b], x_nnode[b], ntip);
                    }
                } catch (const std::exception &e) {
                    errs[th] = e.what();
                }
            });
        for (auto &t : pool) t.join();
        for (auto &e : errs)
            if (!e.empty()) {
                c.err = e;
                return;
            }
        // lanes: the reference rows, then each replicate's rows; outputs: each
        // replicate's ymin (ny values), then every replicate's xmin
        c.lo = ty.plan.lo;
        c.sz = ty.plan.cnt;
        c.depth = ty.walk.depth;
        for (u32 k = 0; k < nb; ++k) c.nx += tx[k].plan.lo.size();
        u64 lane = ny, xo = (u64)nb * ny;
        for (u32 k = 0; k < nb; ++k) {
            const u32 nr = (u32)tx[k].plan.lo.size();
            c.jobs.push_back({c.ops.size(), 0, (u64)k * ny, padded_ops(tx[k].walk), ny});
            emit_walk(tx[k].walk, ty.pos, c.ops);
            c.jobs.push_back({c.ops.size(), lane, xo, padded_ops(ty.walk), nr});
            emit_walk(ty.walk, tx[k].pos, c.ops);
            c.depth = std::max(c.depth, tx[k].walk.depth);
            c.lo.insert(c.lo.end(), tx[k].plan.lo.begin(), tx[k].plan.lo.end());
            c.sz.insert(c.sz.end(), tx[k].plan.cnt.begin(), tx[k].plan.cnt.end());
            for (u32 r : tx[k].plan.cnt) c.xsz.push_back(clade_size(r, ntip));
            lane += nr;
            xo += nr;
        }
    };

    DevBuf dops, dlo, dsz, djobs, dout;
    Chunk cur, next;
    if (nrep) plan_chunk(0, std::min(chunk, nrep), cur);
    u64 xdone = 0;  // replicate rows written to xmin so far
    for (u32 b0 = 0; b0 < nrep; b0 += chunk) {
        CASPHY_REQUIRE(cur.err.empty(), CASPHY_EARG, cur.err);
        const u64 nout = (u64)cur.nb * ny + cur.nx;
        dops.reserve(std::max<size_t>(cur.ops.size(), 1) * 4);
        dlo.reserve(cur.lo.size() * 4);
        dsz.reserve(cur.sz.size() * 4);
        djobs.reserve(cur.jobs.size() * sizeof(WalkJob));
        dout.reserve(std::max<u64>(nout, 1) * 4);
        h2d(dops.p, cur.ops.data(), cur.ops.size() * 4, s);
        h2d(dlo.p, cur.lo.data(), cur.lo.size() * 4, s);
        h2d(dsz.p, cur.sz.data(), cur.sz.size() * 4, s);
        h2d(djobs.p, cur.jobs.data(), cur.jobs.size() * sizeof(WalkJob), s);
        u32 maxl = 0;
        for (const WalkJob &j : cur.jobs) maxl = std::max(maxl, j.nlanes);
        // clades per lane CASPHY_TBE_RPL (1, 2, 4 or 8) and waves per workgroup
        // CASPHY_TBE_WAVES (1 or 4); 2 x 1 measured fastest on 65,536-tip trees
        // (244 jobs/s against 198 for 1 x 1 and 229 for 4 x 4)
        const char *re = std::getenv("CASPHY_TBE_RPL");
        const char *we = std::getenv("CASPHY_TBE_WAVES");
        const u32 rpl = re ? (u32)std::atoi(re) : 2u, waves = we && std::atoi(we) == 4 ? 4u : 1u;
        const size_t lds = (size_t)waves * cur.depth * rpl * 64 * 4;
        const dim3 grid(ceil_div(std::max(maxl, 1u), 64 * waves * rpl), (u32)cur.jobs.size());
        launch_walk(rpl, waves, grid, lds, s, dops.as<u32>(), dlo.as<u32>(), dsz.as<u32>(), djobs.as<WalkJob>(), ntip,
                    init, cur.depth, dout.as<u32>());
        CASPHY_HIP(hipGetLastError());
        // plan the next chunk while this one runs
        const u32 nb0 = b0 + chunk;
        std::thread planner;
        if (nb0 < nrep)
            planner = std::thread([&] {
                try {
                    plan_chunk(nb0, std::min(chunk, nrep - nb0), next);
                } catch (const std::exception &e) {
                    next.err = e.what();
                }
            });
        try {
            d2h(ymin + (u64)b0 * ny, dout.p, (size_t)cur.nb * ny * 4, s);
            if (cur.nx) {
                d2h(xmin + xdone, dout.as<u32>() + (u64)cur.nb * ny, cur.nx * 4, s);
                std::copy(cur.xsz.begin(), cur.xsz.end(), xsize + xdone);
            }
        } catch (...) {
            if (planner.joinable()) planner.join();
            throw;
        }
        if (planner.joinable()) planner.join();
        xdone += cur.nx;
        std::swap(cur, next);
    }
}

}  // namespace casphy
