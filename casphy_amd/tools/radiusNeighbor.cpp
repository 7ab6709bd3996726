// radiusNeighbor <prefix> <radius>
// Drop-in for cu/radiusNeighbor.cu (called by R/CasPhy_functions.R:315,2000,
// 2024,2042).  Reads <prefix>.bmat ([W, n] + word-major u32) and <prefix>.w
// (32*W f64); writes <prefix>.idx (the ascending neighbour lists of cells
// 0..n-1, concatenated) and <prefix>.nNeighbor ([1, n] + list lengths).
#include "tool_common.hpp"

int main(int argc, char **argv)
{
    using namespace tool;
    Timer t;
    if (argc < 3) die("usage: radiusNeighbor <prefix> <radius>");
    const std::string prefix = argv[1];
    const long radius = parse_long(argv[2], "radius");
    BinMat m = read_bin(prefix + ".bmat");
    const uint32_t W = m.d0, n = m.d1;
    if (n == 0 || W == 0) die("empty .bmat");
    std::vector<double> w = read_weights(prefix + ".w", (size_t)W * 32);
    std::vector<uint32_t> counts(n);
    uint32_t *idx = nullptr;
    uint64_t total = 0;
    check(casphy_set_device(device_index()), "casphy_set_device");
    check(casphy_radius_bit(m.data.data(), W, n, weights_or_null(w), (int32_t)radius, counts.data(), &idx, &total),
          "casphy_radius_bit");
    FILE *f = open_out(prefix + ".idx");
    write_all(f, idx, total * 4, prefix + ".idx");
    std::fclose(f);
    casphy_free(idx);
    write_bin(prefix + ".nNeighbor", 1, n, counts.data(), n);
    std::printf("radiusNeighbor: %.2fs elapsed for %u tips.\n", t.s(), n);
    return 0;
}
