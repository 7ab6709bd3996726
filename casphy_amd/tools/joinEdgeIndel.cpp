// joinEdgeIndel <prefix> <wfile> <nState> <minSize> <minRatio>
// Drop-in for cu/joinEdgeIndel.cu (called by R/CasPhy_functions.R:612,2209,
// 2592).  Reads <prefix>.tbmat (states + 1: 0 missing, 1 unmutated, k+1
// allele k), <prefix>.tipSize ([1, n] + n u32) and nState*ncol f64 weights
// (state-major); writes .edge2 and .nodeState2 (states + 1).
#include "tool_common.hpp"

int main(int argc, char **argv)
{
    using namespace tool;
    Timer t;
    if (argc < 6) die("usage: joinEdgeIndel <prefix> <wfile> <nState> <minSize> <minRatio>");
    const std::string prefix = argv[1];
    const long nstate = parse_long(argv[3], "nState");
    const long min_size = parse_long(argv[4], "minSize");
    const float min_ratio = std::strtof(argv[5], nullptr);
    if (nstate < 1 || min_size < 0) die("nState must be >= 1 and minSize >= 0");
    BinMat m = read_bin(prefix + ".tbmat");
    const uint32_t n = m.d0, ncol = m.d1;
    std::vector<double> w = read_weights(argv[2], (size_t)nstate * ncol);
    std::vector<uint32_t> tip(n);
    {
        FILE *f = std::fopen((prefix + ".tipSize").c_str(), "rb");
        std::fseek(f, 8, SEEK_SET);
        if (std::fread(tip.data(), 4, n, f) != n) die(prefix + ".tipSize: fewer than n sizes");
        std::fclose(f);
    }
    if (n == 1) {
        write_bin(prefix + ".edge2", 0, 2, nullptr, 0);
        write_bin(prefix + ".nodeState2", 0, ncol, m.data.data(), ncol);
        return 0;
    }
    std::vector<int32_t> mat((size_t)n * ncol);
    for (size_t i = 0; i < mat.size(); ++i) mat[i] = (int32_t)m.data[i] - 1;
    std::vector<uint32_t> edge((size_t)(2 * n - 2) * 2);
    std::vector<int32_t> state((size_t)(2 * n - 1) * ncol);
    check(casphy_set_device(device_index()), "casphy_set_device");
    check(casphy_join_indel(mat.data(), n, ncol, weights_or_null(w), (uint32_t)nstate, tip.data(),
                            (uint32_t)min_size, min_ratio, edge.data(), state.data(), nullptr),
          "casphy_join_indel");
    std::vector<uint32_t> eo(edge.size()), order(n - 1), rows((size_t)(n - 1) * ncol);
    check(casphy_cladewise(edge.data(), n, eo.data(), order.data()), "casphy_cladewise");
    for (uint32_t k = 0; k < n - 1; ++k)
        for (uint32_t s = 0; s < ncol; ++s)
            rows[(size_t)k * ncol + s] = (uint32_t)(state[(size_t)order[k] * ncol + s] + 1);
    write_bin(prefix + ".edge2", 2 * n - 2, 2, eo.data(), eo.size());
    write_bin(prefix + ".nodeState2", n - 1, ncol, rows.data(), rows.size());
    std::printf("joinEdgeIndel: %.2fs elapsed for %u tips in GPU %d.\n", t.s(), n, device_index());
    return 0;
}
