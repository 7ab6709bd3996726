// tryRadius <prefix> <radius> <nReplicate>
// Drop-in for cu/tryRadius.cu (called by R/CasPhy_functions.R:2017).  Reads
// <prefix>.bmat and <prefix>.w; writes <prefix>.nNeighbor ([2, R] + the R
// query rows + their neighbour counts).  Query rows come from a seeded
// splitmix64 (CASPHY_SEED) instead of the reference's time-seeded rand().
#include "tool_common.hpp"

int main(int argc, char **argv)
{
    using namespace tool;
    Timer t;
    if (argc < 4) die("usage: tryRadius <prefix> <radius> <nReplicate>");
    const std::string prefix = argv[1];
    const long radius = parse_long(argv[2], "radius");
    const long R = parse_long(argv[3], "nReplicate");
    if (R < 0) die("nReplicate must be >= 0");
    BinMat m = read_bin(prefix + ".bmat");
    const uint32_t W = m.d0, n = m.d1;
    if (n == 0 || W == 0) die("empty .bmat");
    std::vector<double> w = read_weights(prefix + ".w", (size_t)W * 32);
    std::vector<uint32_t> q = query_rows(n, (uint32_t)R), out(2 * (size_t)R);
    check(casphy_set_device(device_index()), "casphy_set_device");
    check(casphy_try_radius_bit(m.data.data(), W, n, weights_or_null(w), (int32_t)radius, q.data(), (uint32_t)R,
                                out.data() + R),
          "casphy_try_radius_bit");
    std::copy(q.begin(), q.end(), out.begin());
    write_bin(prefix + ".nNeighbor", 2, (uint32_t)R, out.data(), out.size());
    std::printf("tryRadius: %.2fs elapsed for %u tips.\n", t.s(), n);
    return 0;
}
