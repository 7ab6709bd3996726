// tipMerge <prefix> <p>
// Drop-in for cu/tipMerge.cu (called by R/CasPhy_functions.R:316,2025).  Reads
// <prefix>.nNeighbor ([1, n] + list lengths) and <prefix>.idx (the ascending
// neighbour lists, concatenated); writes <prefix>.member, one category label
// per line (read by R with read.table).
#include "tool_common.hpp"

int main(int argc, char **argv)
{
    using namespace tool;
    Timer t;
    if (argc < 3) die("usage: tipMerge <prefix> <p>");
    const std::string prefix = argv[1];
    char *end = nullptr;
    const double p = std::strtod(argv[2], &end);
    if (end == argv[2] || *end) die("tipMerge: p must be a number");
    BinMat nb = read_bin(prefix + ".nNeighbor");
    const uint32_t n = nb.d1;
    if ((size_t)nb.d0 * nb.d1 != n) die(prefix + ".nNeighbor: expected a [1, n] header");
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; ++i) total += nb.data[i];
    std::vector<uint32_t> idx(total);
    FILE *f = std::fopen((prefix + ".idx").c_str(), "rb");
    if (!f) die("cannot open input " + prefix + ".idx");
    if (total && std::fread(idx.data(), 4, total, f) != total) die(prefix + ".idx: shorter than the list lengths");
    std::fclose(f);
    std::vector<int32_t> member(n);
    check(casphy_set_device(device_index()), "casphy_set_device");
    check(casphy_tip_merge(nb.data.data(), idx.data(), n, p, member.data()), "casphy_tip_merge");
    FILE *o = open_out(prefix + ".member");
    for (uint32_t i = 0; i < n; ++i)
        if (std::fprintf(o, "%d\n", member[i]) < 0) die("write failed: " + prefix + ".member");
    if (std::fclose(o) != 0) die("close failed: " + prefix + ".member");
    std::printf("tipMerge: %.2fs elapsed for %u tips.\n", t.s(), n);
    return 0;
}
