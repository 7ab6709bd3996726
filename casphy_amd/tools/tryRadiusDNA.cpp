// tryRadiusDNA <prefix> <radius> <nReplicate>
// Drop-in for cu/tryRadiusDNA.cu (called by R/CasPhy_functions.R:2014).
#include "tool_common.hpp"

int main(int argc, char **argv)
{
    using namespace tool;
    Timer t;
    if (argc < 4) die("usage: tryRadiusDNA <prefix> <radius> <nReplicate>");
    const std::string prefix = argv[1];
    const long radius = parse_long(argv[2], "radius");
    const long R = parse_long(argv[3], "nReplicate");
    if (R < 0) die("nReplicate must be >= 0");
    BinMat m = read_bin(prefix + ".tbmat");
    const uint32_t n = m.d0, nw = m.d1;
    if (n == 0 || nw == 0) die("empty .tbmat");
    std::vector<uint32_t> q = query_rows(n, (uint32_t)R), out(2 * (size_t)R);
    check(casphy_set_device(device_index()), "casphy_set_device");
    check(casphy_try_radius_dna(m.data.data(), n, nw, (int32_t)radius, q.data(), (uint32_t)R, out.data() + R),
          "casphy_try_radius_dna");
    std::copy(q.begin(), q.end(), out.begin());
    double avg = 0;
    for (long i = 0; i < R; ++i) avg += out[R + i];
    std::printf("Average nNeighbor: %.2f\n", R ? avg / R : 0.0);
    write_bin(prefix + ".nNeighbor", 2, (uint32_t)R, out.data(), out.size());
    std::printf("tryRadiusDNA: %.2fs elapsed for %u tips.\n", t.s(), n);
    return 0;
}
