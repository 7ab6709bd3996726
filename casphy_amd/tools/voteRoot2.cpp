// voteRoot2 <matfile> <prefix>
// Drop-in for cu/voteRoot2.cu (called by R/DNA_functions.R:130): reads
// <prefix>.nodeState ([nnode, nbin] + node rows of byte codes), <prefix>.edge
// ([nedge, 2] + 1-based (parent, child) pairs), <prefix>.outmatIdx ([1, n] + n
// 0-based row indices) and those rows of <matfile> (a .bin matrix: 2 header
// words, then rows of nbin words); writes <prefix>.rootVote ([2, n] + the n
// 0-based voted nodes + the n shared-position counts).
#include "tool_common.hpp"

int main(int argc, char **argv)
{
    using namespace tool;
    Timer t;
    if (argc < 3) die("usage: voteRoot2 <matfile> <prefix>");
    const std::string matfile = argv[1], prefix = argv[2];
    BinMat st = read_bin(prefix + ".nodeState");
    BinMat ed = read_bin(prefix + ".edge", false);
    const uint32_t nnode = st.d0, nbin = st.d1, nedge = ed.d0;
    std::vector<uint32_t> edge((size_t)nedge * 2);
    std::vector<uint32_t> idx;
    {
        FILE *f = std::fopen((prefix + ".edge").c_str(), "rb");
        if (!f) die("cannot open input " + prefix + ".edge");
        std::fseek(f, 8, SEEK_SET);
        if (std::fread(edge.data(), 4, edge.size(), f) != edge.size()) die(prefix + ".edge: truncated payload");
        std::fclose(f);
        BinMat oi = read_bin(prefix + ".outmatIdx", false);  // [1, n]: the count is the second word
        idx.resize(oi.d1);
        f = std::fopen((prefix + ".outmatIdx").c_str(), "rb");
        if (!f) die("cannot open input " + prefix + ".outmatIdx");
        std::fseek(f, 8, SEEK_SET);
        if (std::fread(idx.data(), 4, idx.size(), f) != idx.size()) die(prefix + ".outmatIdx: truncated payload");
        std::fclose(f);
    }
    const uint32_t n = (uint32_t)idx.size();
    std::vector<uint32_t> rows((size_t)n * nbin), iota(n);
    {
        FILE *f = std::fopen(matfile.c_str(), "rb");
        if (!f) die("cannot open input " + matfile);
        for (uint32_t k = 0; k < n; ++k) {
            iota[k] = k;
            if (!nbin) continue;
            const long long off = 8 + 4LL * nbin * idx[k];
            if (std::fseek(f, (long)off, SEEK_SET) != 0 || std::fread(&rows[(size_t)k * nbin], 4, nbin, f) != nbin)
                die(matfile + ": row " + std::to_string(idx[k]) + " is beyond the file");
        }
        std::fclose(f);
    }
    std::vector<uint32_t> out((size_t)2 * n + 1);
    check(casphy_set_device(device_index()), "casphy_set_device");
    check(casphy_vote_root(st.data.data(), nnode, nbin, edge.data(), nedge, rows.data(), n, iota.data(), n,
                           out.data(), out.data() + n),
          "casphy_vote_root");
    write_bin(prefix + ".rootVote", 2, n, out.data(), (size_t)2 * n);
    std::printf("voteRoot: %.2fs elapsed for %u nodes and %u outgroups.\n", t.s(), nnode, n);
    return 0;
}
