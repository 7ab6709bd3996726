// tbeSupport <reference tree> <bootstrap trees> <output prefix>
// Transfer bootstrap support of a reference tree's clades from Newick files:
// the reference is the first tree of its file, the replicates every tree of
// the second (tips matched by label).  The clade bitsets of all trees are
// built on the device and scored there (casphy_transfer_trees, the batched
// form of TransferDisStream, cu/TransferDisStream.cu), then normalised
// (casphy_tbe_support).  Writes <prefix>_tbe.txt: "node,size,support", one
// line per clade row in nodeDesBig's order (src/functions.cpp:1277-1366),
// node = the clade's R node id, support NA for splits with one tip on a side.
// No counterpart executable in the reference (R compares trees with
// approximateCompare2_AVX2, R/bootstrap_functions.R:34-84).
#include <cmath>
#include <string>

#include "tool_common.hpp"

namespace {

std::string slurp(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) tool::die("cannot open input " + path);
    std::string s;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, got);
    std::fclose(f);
    return s;
}

struct Tree {
    uint32_t ntip = 0, nnode = 0, nedge = 0;
    std::vector<int32_t> edge;  // parent column, then child column
    std::string labels;         // NUL-separated
};

// Every tree of `text`, tips numbered by `order` when given.
std::vector<Tree> read_all(const std::string &text, const std::string *order, size_t limit)
{
    std::vector<Tree> out;
    uint64_t pos = 0;
    while (pos < text.size() && out.size() < limit) {
        uint64_t used = 0, nlab = 0;
        Tree t;
        int32_t *e = nullptr;
        char *lab = nullptr;
        tool::check(casphy_newick_read(text.data() + pos, text.size() - pos, order ? order->data() : nullptr,
                                       order ? order->size() : 0, &used, &t.ntip, &t.nnode, &e, &t.nedge, &lab,
                                       &nlab),
                    "casphy_newick_read");
        pos += used;
        if (!t.ntip) break;
        t.edge.assign(e, e + 2 * (size_t)t.nedge);
        t.labels.assign(lab, nlab);
        casphy_free(e);
        casphy_free(lab);
        out.push_back(std::move(t));
    }
    return out;
}

}  // namespace

int main(int argc, char **argv)
{
    using namespace tool;
    Timer tm;
    if (argc < 4) die("usage: tbeSupport <reference tree> <bootstrap trees> <output prefix>");
    const std::string prefix = argv[3];
    std::vector<Tree> ref = read_all(slurp(argv[1]), nullptr, 1);
    if (ref.empty()) die(std::string(argv[1]) + ": no tree");
    const Tree &y = ref[0];
    if (y.nnode < 2) die(std::string(argv[1]) + ": the reference tree needs a non-root internal node");
    std::vector<Tree> reps = read_all(slurp(argv[2]), &y.labels, (size_t)-1);
    const uint32_t nrep = (uint32_t)reps.size(), ny = y.nnode - 1, ntip = y.ntip;
    std::vector<int32_t> xa, xd;
    std::vector<uint32_t> eoff(nrep + 1, 0), xn(nrep);
    size_t xrows = 0;
    for (uint32_t b = 0; b < nrep; ++b) {
        const Tree &t = reps[b];
        xa.insert(xa.end(), t.edge.begin(), t.edge.begin() + t.nedge);
        xd.insert(xd.end(), t.edge.begin() + t.nedge, t.edge.end());
        eoff[b + 1] = eoff[b] + t.nedge;
        xn[b] = t.nnode;
        xrows += t.nnode ? t.nnode - 1 : 0;
    }
    std::vector<uint32_t> ymin((size_t)nrep * ny + 1), ysize(ny), xmin(xrows + 1), xsize(xrows + 1);
    check(casphy_set_device(device_index()), "casphy_set_device");
    check(casphy_transfer_trees(y.edge.data(), y.edge.data() + y.nedge, y.nedge, y.nnode, xa.data(), xd.data(),
                                eoff.data(), xn.data(), nrep, ntip, ymin.data(), ysize.data(), xmin.data(),
                                xsize.data()),
          "casphy_transfer_trees");
    std::vector<double> sup(ny);
    check(casphy_tbe_support(ymin.data(), nrep, ny, ysize.data(), ntip, sup.data()), "casphy_tbe_support");
    // clade rows: the internal children in backward edge order
    std::vector<int32_t> node;
    for (uint32_t e = y.nedge; e-- > 0;)
        if ((uint32_t)y.edge[y.nedge + e] > ntip) node.push_back(y.edge[y.nedge + e]);
    if (node.size() != ny) die("internal: clade rows do not match the reference tree");
    const std::string path = prefix + "_tbe.txt";
    FILE *o = open_out(path);
    std::string body = "node,size,support\n";
    char line[96];
    for (uint32_t k = 0; k < ny; ++k) {
        if (std::isnan(sup[k]))
            std::snprintf(line, sizeof line, "%d,%u,NA\n", node[k], ysize[k]);
        else
            std::snprintf(line, sizeof line, "%d,%u,%.17g\n", node[k], ysize[k], sup[k]);
        body += line;
    }
    write_all(o, body.data(), body.size(), path);
    if (std::fclose(o) != 0) die("close failed: " + path);
    std::printf("tbeSupport: %.2fs for %u clades x %u replicates on %u tips.\n", tm.s(), ny, nrep, ntip);
    return 0;
}
