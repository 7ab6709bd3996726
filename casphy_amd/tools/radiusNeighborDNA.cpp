// radiusNeighborDNA <prefix> <radius>
// Drop-in for cu/radiusNeighborDNA.cu (called by R/CasPhy_functions.R:1887,
// 2000,2024,2042).  Reads <prefix>.tbmat (fasta2bin .bin: [n, nwRaw] + 4
// byte codes per u32); writes <prefix>.idx and <prefix>.nNeighbor.
#include "tool_common.hpp"

int main(int argc, char **argv)
{
    using namespace tool;
    Timer t;
    if (argc < 3) die("usage: radiusNeighborDNA <prefix> <radius>");
    const std::string prefix = argv[1];
    const long radius = parse_long(argv[2], "radius");
    BinMat m = read_bin(prefix + ".tbmat");
    const uint32_t n = m.d0, nw = m.d1;
    if (n == 0 || nw == 0) die("empty .tbmat");
    std::vector<uint32_t> counts(n);
    uint32_t *idx = nullptr;
    uint64_t total = 0;
    check(casphy_set_device(device_index()), "casphy_set_device");
    check(casphy_radius_dna(m.data.data(), n, nw, (int32_t)radius, counts.data(), &idx, &total), "casphy_radius_dna");
    FILE *f = open_out(prefix + ".idx");
    write_all(f, idx, total * 4, prefix + ".idx");
    std::fclose(f);
    casphy_free(idx);
    write_bin(prefix + ".nNeighbor", 1, n, counts.data(), n);
    std::printf("radiusNeighborDNA: %.2fs elapsed for %u tips.\n", t.s(), n);
    return 0;
}
