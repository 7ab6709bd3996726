// node_keep_tool.hpp -- the shared main of the nodeKeepBit / nodeKeepIndel
// executables (cu/nodeKeepBit.cu:75-227, cu/nodeKeepIndel.cu:50-169): reads
// <prefix>.nodeState2 ([nnode, width] + rows), <prefix>.edge2 ([nedge, 2] +
// 0-based edges) and <prefix>.tboutmat ([nout, ...] + nout rows of `width`
// words); writes <prefix>.retainNode, one kept internal node id per line.
#pragma once

#include "tool_common.hpp"

namespace tool {

inline int node_keep_main(int argc, char **argv, bool indel)
{
    const char *name = indel ? "nodeKeepIndel" : "nodeKeepBit";
    Timer t;
    if (argc < 3) die(std::string("usage: ") + name + " <prefix> <cutoff>");
    const std::string prefix = argv[1];
    const long cutoff = parse_long(argv[2], "cutoff");
    if (cutoff < 0 || cutoff > 0xFFFFFFFFL) die("cutoff out of range");
    BinMat st = read_bin(prefix + ".nodeState2");
    BinMat ed = read_bin(prefix + ".edge2");
    if (ed.d1 != 2) die(prefix + ".edge2: expected a [nedge, 2] header");
    const uint32_t nnode = st.d0, width = st.d1;
    BinMat om = read_bin(prefix + ".tboutmat", false);
    const uint32_t nout = om.d0;
    om.data.resize((size_t)nout * width);
    {
        FILE *f = std::fopen((prefix + ".tboutmat").c_str(), "rb");
        if (!f) die("cannot open input " + prefix + ".tboutmat");
        std::fseek(f, 8, SEEK_SET);
        if (std::fread(om.data.data(), 4, om.data.size(), f) != om.data.size())
            die(prefix + ".tboutmat: truncated payload");
        std::fclose(f);
    }
    std::vector<uint32_t> kept(nnode ? nnode : 1);
    uint32_t nk = 0;
    check(casphy_set_device(device_index()), "casphy_set_device");
    if (indel)
        check(casphy_node_keep_indel(ed.data.data(), ed.d0, st.data.data(), nnode, width, om.data.data(), nout,
                                     (uint32_t)cutoff, kept.data(), &nk),
              "casphy_node_keep_indel");
    else
        check(casphy_node_keep_bit(ed.data.data(), ed.d0, st.data.data(), nnode, width, om.data.data(), nout,
                                   (uint32_t)cutoff, kept.data(), &nk),
              "casphy_node_keep_bit");
    FILE *o = open_out(prefix + ".retainNode");
    for (uint32_t i = 0; i < nk; ++i)
        if (std::fprintf(o, "%u\n", kept[i]) < 0) die("write failed: " + prefix + ".retainNode");
    if (std::fclose(o) != 0) die("close failed: " + prefix + ".retainNode");
    std::printf("%s: %.2fs elapsed for %u nodes and %u outgroups.\n", name, t.s(), nnode, nout);
    return 0;
}

}  // namespace tool
