// joinEdgeBitReverse <prefix> <wfile> <ancestralType 0|1>
// Drop-in for cu/joinEdgeBitReverse.cu (called by R/CasPhy_functions.R:607,
// 2204, 2587).  Reads <prefix>.tbmat ([n, W] + n*W u32, cell-major) and the
// 32*W f64 weights; writes <prefix>.edge2 ([2n-2, 2] + cladewise relabelled
// edges) and <prefix>.nodeState2 ([n-1, W] + the root state, then each
// internal node's seq & ~bistate in relabel order).
#include "tool_common.hpp"

int main(int argc, char **argv)
{
    using namespace tool;
    Timer t;
    if (argc < 4) die("usage: joinEdgeBitReverse <prefix> <wfile> <ancestralType>");
    const std::string prefix = argv[1];
    const int anc = (int)parse_long(argv[3], "ancestralType");
    BinMat m = read_bin(prefix + ".tbmat");
    const uint32_t n = m.d0, W = m.d1;
    std::vector<double> w = read_weights(argv[2], (size_t)W * 32);
    if (n == 1) {  // the reference emits no edge and the lone tip as the root state
        write_bin(prefix + ".edge2", 0, 2, nullptr, 0);
        write_bin(prefix + ".nodeState2", 0, W, m.data.data(), W);
        return 0;
    }
    std::vector<uint32_t> edge((size_t)(2 * n - 2) * 2), state((size_t)(2 * n - 1) * W);
    check(casphy_set_device(device_index()), "casphy_set_device");
    check(casphy_join_bit(m.data.data(), n, W, weights_or_null(w), anc, edge.data(), state.data(), nullptr),
          "casphy_join_bit");
    std::vector<uint32_t> eo(edge.size()), order(n - 1), rows((size_t)(n - 1) * W);
    check(casphy_cladewise(edge.data(), n, eo.data(), order.data()), "casphy_cladewise");
    for (uint32_t k = 0; k < n - 1; ++k)
        std::copy(state.begin() + (size_t)order[k] * W, state.begin() + (size_t)(order[k] + 1) * W,
                  rows.begin() + (size_t)k * W);
    write_bin(prefix + ".edge2", 2 * n - 2, 2, eo.data(), eo.size());
    write_bin(prefix + ".nodeState2", n - 1, W, rows.data(), rows.size());
    std::printf("joinEdgeBit: %.2fs elapsed for %u tips in GPU %d.\n", t.s(), n, device_index());
    return 0;
}
