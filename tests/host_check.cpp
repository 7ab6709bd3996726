// host_check.cpp -- CPU driver of the library's host-only planning code
// (casphy_amd/csrc/plan.cpp, newick.cpp), built by tests/test_host_plan.py
// with the address and undefined-behaviour sanitizers.  It also replays the
// transfer-distance walk programs with a scalar loop that follows the device
// kernel's op semantics, so the walks are pinned against the oracle on the CPU.
//
//   host_check walk <dir>      dir/meta.txt: "ntip y_nedge y_nnode nrep" then
//                              per replicate "nedge nnode"; dir/y_anc.bin,
//                              y_des.bin, x_anc.bin, x_des.bin (int32, 1-based)
//                              -> dir/ymin.bin (nrep x ny), ysize.bin,
//                                 xmin.bin, xsize.bin, depth.txt
//   host_check clades <dir>    meta "ntip nedge nnode"; anc.bin, des.bin
//                              -> order.bin, lo.bin, cnt.bin, size.bin
//   host_check nodekeep <dir>  meta "nedge ntip"; edge.bin (u32 pairs),
//                              retain.bin (u32, ntip - 1) -> kept.bin
//   host_check newick <dir>    text.nwk -> per tree: trees.txt lines
//                              "ntip nnode nedge", anc/des appended to
//                              anc.bin / des.bin
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../casphy_amd/csrc/newick.hpp"
#include "../casphy_amd/csrc/plan.hpp"

using namespace casphy;

template <class T> std::vector<T> read_bin(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot read " + path);
    f.seekg(0, std::ios::end);
    const size_t n = (size_t)f.tellg() / sizeof(T);
    f.seekg(0);
    std::vector<T> v(n);
    if (n) f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(n * sizeof(T)));
    return v;
}

template <class T> void write_bin(const std::string &path, const std::vector<T> &v)
{
    std::ofstream f(path, std::ios::binary);
    if (!v.empty()) f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

// The device kernel's loop for one lane (its clade run [lo, lo + sz)).
u32 replay(const std::vector<u32> &ops, size_t off, size_t nops, u32 lo, u32 sz, u32 ntip, u32 init, u32 depth)
{
    std::vector<u32> stk(depth + 1, 0xDEADBEEFu);
    u32 top = 0, best = init, sp = 0;
    for (size_t i = 0; i < nops; ++i) {
        const u32 w = ops[off + i], arg = w & ARG_MASK;
        if (w & OP_POS) {
            const u32 hit = (arg - lo) < sz ? 1u : 0u;
            if (w & OP_STACK) {
                if (sp >= depth) throw std::runtime_error("walk stack overflow");
                stk[sp++] = top;
                top = hit;
            } else {
                top += hit;
            }
        } else {
            if (w & OP_STACK) {
                if (sp == 0) throw std::runtime_error("walk stack underflow");
                top += stk[--sp];
            }
            if (w & OP_ROW) {
                const u32 h = sz + arg - 2u * top;
                const u32 hc = ntip - h;
                const u32 d = h < hc ? h : hc;
                best = d < best ? d : best;
            }
        }
    }
    return best;
}

int walk(const std::string &dir)
{
    std::ifstream meta(dir + "/meta.txt");
    u32 ntip, ynedge, ynnode, nrep;
    meta >> ntip >> ynedge >> ynnode >> nrep;
    std::vector<u32> xne(nrep), xnn(nrep);
    for (u32 b = 0; b < nrep; ++b) meta >> xne[b] >> xnn[b];
    const auto yanc = read_bin<int32_t>(dir + "/y_anc.bin"), ydes = read_bin<int32_t>(dir + "/y_des.bin");
    const auto xanc = read_bin<int32_t>(dir + "/x_anc.bin"), xdes = read_bin<int32_t>(dir + "/x_des.bin");
    const TreeData ty = tree_data(yanc.data(), ydes.data(), ynedge, ynnode, ntip);
    const u32 ny = (u32)ty.plan.lo.size(), init = 32u * ty.plan.nbin;
    std::vector<u32> ymin, ysize(ny), xmin, xsize;
    for (u32 y = 0; y < ny; ++y) ysize[y] = std::min(ty.plan.cnt[y], ntip - ty.plan.cnt[y]);
    u32 depth = ty.walk.depth;
    size_t eo = 0;
    for (u32 b = 0; b < nrep; ++b) {
        const TreeData tx = tree_data(xanc.data() + eo, xdes.data() + eo, xne[b], xnn[b], ntip);
        eo += xne[b];
        depth = std::max(depth, tx.walk.depth);
        std::vector<u32> ops;
        emit_walk(tx.walk, ty.pos, ops);
        const size_t na = ops.size();
        if (na != padded_ops(tx.walk) || na % OP_BLOCK) throw std::runtime_error("walk not padded");
        emit_walk(ty.walk, tx.pos, ops);
        for (u32 y = 0; y < ny; ++y)
            ymin.push_back(replay(ops, 0, na, ty.plan.lo[y], ty.plan.cnt[y], ntip, init, std::max(ty.walk.depth,
                                                                                                   tx.walk.depth)));
        for (size_t x = 0; x < tx.plan.lo.size(); ++x) {
            xmin.push_back(replay(ops, na, ops.size() - na, tx.plan.lo[x], tx.plan.cnt[x], ntip, init,
                                  std::max(ty.walk.depth, tx.walk.depth)));
            xsize.push_back(std::min(tx.plan.cnt[x], ntip - tx.plan.cnt[x]));
        }
    }
    write_bin(dir + "/ymin.bin", ymin);
    write_bin(dir + "/ysize.bin", ysize);
    write_bin(dir + "/xmin.bin", xmin);
    write_bin(dir + "/xsize.bin", xsize);
    std::ofstream(dir + "/depth.txt") << depth << "\n";
    return 0;
}

int clades(const std::string &dir)
{
    std::ifstream meta(dir + "/meta.txt");
    u32 ntip, nedge, nnode;
    meta >> ntip >> nedge >> nnode;
    const auto anc = read_bin<int32_t>(dir + "/anc.bin"), des = read_bin<int32_t>(dir + "/des.bin");
    const CladePlan p = plan_clades(anc.data(), des.data(), nedge, nnode, ntip);
    write_bin(dir + "/order.bin", p.order);
    write_bin(dir + "/lo.bin", p.lo);
    write_bin(dir + "/cnt.bin", p.cnt);
    write_bin(dir + "/size.bin", p.size);
    return 0;
}

int nodekeep(const std::string &dir)
{
    std::ifstream meta(dir + "/meta.txt");
    u32 nedge, ntip;
    meta >> nedge >> ntip;
    const auto edge = read_bin<u32>(dir + "/edge.bin"), retain = read_bin<u32>(dir + "/retain.bin");
    std::vector<u32> kept(ntip);
    kept.resize(node_keep_walk(edge.data(), nedge, retain.data(), ntip, kept.data()));
    write_bin(dir + "/kept.bin", kept);
    return 0;
}

int newick(const std::string &dir)
{
    std::ifstream f(dir + "/text.nwk", std::ios::binary);
    const std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::ofstream trees(dir + "/trees.txt");
    std::vector<int32_t> anc, des;
    u64 pos = 0;
    for (;;) {
        const NewickTree t = read_newick(text.data(), text.size(), pos);
        if (t.ntip == 0) break;
        trees << t.ntip << " " << t.nnode << " " << t.anc.size() << "\n";
        anc.insert(anc.end(), t.anc.begin(), t.anc.end());
        des.insert(des.end(), t.des.begin(), t.des.end());
    }
    write_bin(dir + "/anc.bin", anc);
    write_bin(dir + "/des.bin", des);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: host_check walk|clades|nodekeep|newick <dir>\n");
        return 2;
    }
    try {
        const std::string cmd = argv[1], dir = argv[2];
        if (cmd == "walk") return walk(dir);
        if (cmd == "clades") return clades(dir);
        if (cmd == "nodekeep") return nodekeep(dir);
        if (cmd == "newick") return newick(dir);
        std::fprintf(stderr, "unknown command %s\n", argv[1]);
        return 2;
    } catch (const Error &e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 3;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 4;
    }
}
