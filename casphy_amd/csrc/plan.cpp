// plan.cpp -- host-side planning of the tree kernels (clade rows, transfer
// walks, nodeKeep's skip rule).  Plain C++: no HIP header.
#include "plan.hpp"

#include <algorithm>
#include <utility>

namespace casphy {

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

TreeShape tree_shape(const int32_t *anc, const int32_t *des, u32 nedge, u32 nnode, u32 ntip)
{
    TreeShape t;
    t.ntip = ntip;
    t.nnode = nnode;
    const u32 nall = ntip + nnode;
    t.kid_off.assign(nall + 1, 0);
    std::vector<unsigned char> child(nall, 0);
    for (u32 e = 0; e < nedge; ++e) {
        ++t.kid_off[(u32)anc[e]];  // anc is 1-based: node anc-1, counted at slot anc
        child[(u32)des[e] - 1] = 1;
    }
    for (u32 v = 0; v < nall; ++v) t.kid_off[v + 1] += t.kid_off[v];
    t.kid.resize(nedge);
    std::vector<u32> fill(t.kid_off.begin(), t.kid_off.end() - 1);
    for (u32 e = 0; e < nedge; ++e) t.kid[fill[(u32)anc[e] - 1]++] = (u32)des[e] - 1;
    t.root = ntip;
    for (u32 v = ntip; v < nall; ++v)
        if (!child[v]) t.root = v;
    // tips below each node: the edges run parents before children (cladewise),
    // so a backward pass finishes every child before its parent
    t.tips.assign(nall, 0);
    for (u32 v = 0; v < ntip; ++v) t.tips[v] = 1;
    for (u32 e = nedge; e-- > 0;) t.tips[(u32)anc[e] - 1] += t.tips[(u32)des[e] - 1];
    t.is_row.assign(nall, 0);
    for (u32 v = ntip; v < nall; ++v) t.is_row[v] = v != t.root;
    return t;
}

WalkSkeleton walk_skeleton(const TreeShape &t)
{
    WalkSkeleton w;
    const u32 nall = t.ntip + t.nnode;
    // heavy child of every internal node (first of the largest)
    std::vector<u32> heavy(nall, ~0u);
    for (u32 v = t.ntip; v < nall; ++v) {
        u32 best = ~0u, bt = 0;
        for (u32 k = t.kid_off[v]; k < t.kid_off[v + 1]; ++k)
            if (best == ~0u || t.tips[t.kid[k]] > bt) {
                best = t.kid[k];
                bt = t.tips[best];
            }
        heavy[v] = best;
    }
    struct Frame {
        u32 v, k;    // node, next child slot
        bool light;  // v is a light child of the frame below
    };
    std::vector<Frame> fr;
    u32 dsp = 0;
    // descend from v along heavy children to a tip
    auto descend = [&](u32 v, bool light) {
        while (v >= t.ntip) {
            CASPHY_REQUIRE(heavy[v] != ~0u, CASPHY_EARG, "transfer walk: internal node without children");
            fr.push_back({v, t.kid_off[v], light});
            light = false;
            v = heavy[v];
        }
        w.ops.push_back(op_tip(v));
        w.depth = std::max(w.depth, ++dsp);
    };
    if (t.root >= t.ntip) descend(t.root, false);
    while (!fr.empty()) {
        Frame &f = fr.back();
        while (f.k < t.kid_off[f.v + 1] && t.kid[f.k] == heavy[f.v]) ++f.k;  // the heavy child is done
        if (f.k < t.kid_off[f.v + 1]) {
            const u32 c = t.kid[f.k++];
            if (c < t.ntip) {
                w.ops.push_back(op_acc(c));
            } else {
                descend(c, true);  // invalidates f
            }
            continue;
        }
        // node v is done: the running value is its count; a light child is
        // then added into its parent's running value
        const Frame done = f;
        fr.pop_back();
        const bool row = t.is_row[done.v] != 0;
        const u32 size = t.tips[done.v];
        CASPHY_REQUIRE(size <= ARG_MASK, CASPHY_EARG, "transfer walk: more than 2^29 - 1 tips");
        if (row) w.ops.push_back(op_row(size));
        if (done.light) {
            w.ops.push_back(op_add());
            --dsp;
        }
    }
    // an ADD that completes a node followed by that node's ROW is one ADD
    // carrying the row (the kernel adds, then evaluates)
    std::vector<u32> merged;
    merged.reserve(w.ops.size());
    for (size_t k = 0; k < w.ops.size(); ++k) {
        const u32 x = w.ops[k];
        if (x == op_add() && k + 1 < w.ops.size() && (w.ops[k + 1] & ~ARG_MASK) == OP_ROW) {
            merged.push_back(op_add() | w.ops[k + 1]);
            ++k;
        } else {
            merged.push_back(x);
        }
    }
    w.ops.swap(merged);
    CASPHY_REQUIRE(w.depth <= WALK_STACK_MAX, CASPHY_EINTERNAL, "transfer walk: stack deeper than expected");
    return w;
}

// Appends a walk with tip ids mapped to positions of the lane tree, padded to
// a multiple of OP_BLOCK words with ACC ops of a position no run contains.
void emit_walk(const WalkSkeleton &w, const std::vector<u32> &pos, std::vector<u32> &ops)
{
    const size_t o = ops.size();
    const size_t n = (w.ops.size() + OP_BLOCK - 1) / OP_BLOCK * OP_BLOCK;
    ops.resize(o + n, op_acc(NOT_IN_TREE));
    for (size_t k = 0; k < w.ops.size(); ++k) {
        const u32 x = w.ops[k];
        ops[o + k] = (x & OP_POS) ? (x & ~ARG_MASK) | pos[x & ARG_MASK] : x;
    }
}

TreeData tree_data(const int32_t *anc, const int32_t *des, u32 nedge, u32 nnode, u32 ntip)
{
    TreeData d;
    d.plan = plan_clades(anc, des, nedge, nnode, ntip);  // validates the edge matrix
    d.pos.assign(ntip, NOT_IN_TREE);
    for (u32 k = 0; k < (u32)d.plan.order.size(); ++k) d.pos[d.plan.order[k]] = k;
    d.walk = walk_skeleton(tree_shape(anc, des, nedge, nnode, ntip));
    return d;
}

u32 node_keep_walk(const u32 *edge, u32 nedge, const u32 *retain, u32 ntip, u32 *kept)
{
    u32 current = ntip - 1, flag = 0, nk = 0;
    for (u32 i = 0; i < nedge; ++i) {
        const u32 a = edge[2 * i], d = edge[2 * i + 1];
        if (d + 1 > ntip) {
            if (flag == 1 && current <= a) continue;
            flag = retain[d - ntip];
            if (flag == 1) {
                current = d;
                kept[nk++] = d;
            }
        }
    }
    return nk;
}

}  // namespace casphy
