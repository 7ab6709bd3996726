// TransferDisStream <Yfile> <Xfile> <outprefix> <ntip>
// Drop-in for cu/TransferDisStream.cu.  Reads two bitset files ([nrow, ncol]
// + nrow*ncol u32; the X header's ncol is ignored as in the reference, :147)
// and writes <outprefix>_ymin.bin ([2, ny] + ymin + ysize) and
// <outprefix>_xmin.bin ([2, nx] + xmin + xsize).
#include "tool_common.hpp"

int main(int argc, char **argv)
{
    using namespace tool;
    if (argc < 5) die("usage: TransferDisStream <Yfile> <Xfile> <outprefix> <ntip>");
    const long ntip = parse_long(argv[4], "ntip");
    BinMat y = read_bin(argv[1]);
    BinMat xh = read_bin(argv[2], false);
    const uint32_t ny = y.d0, ncol = y.d1, nx = xh.d0;
    std::vector<uint32_t> x((size_t)nx * ncol);
    {
        FILE *f = std::fopen(argv[2], "rb");
        std::fseek(f, 8, SEEK_SET);
        if (std::fread(x.data(), 4, x.size(), f) != x.size()) die(std::string(argv[2]) + ": truncated payload");
        std::fclose(f);
    }
    std::printf("calculate crossXOR between 1st mat with %d rows and 2nd mat with %d rows.\n", ny, nx);
    Timer t;
    std::vector<uint32_t> yo(2 * (size_t)ny), xo(2 * (size_t)nx);
    check(casphy_set_device(device_index()), "casphy_set_device");
    check(casphy_transfer(y.data.data(), ny, x.data(), nx, ncol, (uint32_t)ntip, yo.data(), yo.data() + ny, xo.data(),
                          xo.data() + nx),
          "casphy_transfer");
    std::printf("Running time on GPU: %.5fs\n", t.s());
    const std::string p = argv[3];
    write_bin(p + "_ymin.bin", 2, ny, yo.data(), yo.size());
    write_bin(p + "_xmin.bin", 2, nx, xo.data(), xo.size());
    return 0;
}
