// tool_common.hpp -- file I/O shared by the drop-in executables.
//
// Formats (unchanged from the reference): a "[u32 dim0, u32 dim1] + payload"
// binary matrix as written by write.mat2bin (R/CasPhy_functions.R:1551-1583),
// headerless f64 weight files (writeBin(as.numeric(w))) and headerless u32
// .idx lists.  Failures print to stderr and exit non-zero (the reference
// printed and exited 0; R only checks whether output files exist).
#pragma once

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "casphy_hip.h"

namespace tool {

[[noreturn]] inline void die(const std::string &msg)
{
    std::fprintf(stderr, "%s\n", msg.c_str());
    std::exit(1);
}

inline void check(int rc, const char *what)
{
    if (rc != CASPHY_OK) die(std::string(what) + " failed (" + std::to_string(rc) + "): " + casphy_last_error());
}

struct BinMat {
    uint32_t d0 = 0, d1 = 0;
    std::vector<uint32_t> data;
};

inline BinMat read_bin(const std::string &path, bool need_payload = true)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) die("cannot open input " + path);
    BinMat m;
    uint32_t dim[2];
    if (std::fread(dim, 4, 2, f) != 2) die(path + ": truncated header");
    m.d0 = dim[0];
    m.d1 = dim[1];
    if (need_payload) {
        const size_t cnt = (size_t)m.d0 * m.d1;
        m.data.resize(cnt);
        if (std::fread(m.data.data(), 4, cnt, f) != cnt) die(path + ": truncated payload");
    }
    std::fclose(f);
    return m;
}

inline FILE *open_out(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) die("cannot create output " + path);
    return f;
}

inline void write_all(FILE *f, const void *p, size_t bytes, const std::string &path)
{
    if (bytes && std::fwrite(p, 1, bytes, f) != bytes) die("write failed: " + path);
}

inline void write_bin(const std::string &path, uint32_t d0, uint32_t d1, const uint32_t *data, size_t count)
{
    FILE *f = open_out(path);
    const uint32_t dim[2] = {d0, d1};
    write_all(f, dim, 8, path);
    write_all(f, data, count * 4, path);
    if (std::fclose(f) != 0) die("close failed: " + path);
}

inline std::vector<double> read_weights(const std::string &path, size_t count)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) die("cannot open weight file " + path);
    std::vector<double> w(count);
    if (std::fread(w.data(), 8, count, f) != count) die(path + ": fewer than " + std::to_string(count) + " weights");
    std::fclose(f);
    return w;
}

// Unit weights need no plane decomposition: pass NULL then.
inline const double *weights_or_null(const std::vector<double> &w)
{
    for (double v : w)
        if (v != 1.0) return w.data();
    return nullptr;
}

// Integer arguments are read as the reference reads them, with atoi
// (cu/radiusNeighbor.cu:177, cu/tryRadius.cu:61-62, cu/joinEdgeIndel.cu:254-257,
// cu/TransferDisStream.cu:138): leading blanks and a sign, then the digits up
// to the first other character; no digits read as 0.
inline long parse_long(const char *s, const char *what)
{
    (void)what;
    return (long)std::atoi(s);
}

// splitmix64: explicit, seeded replacement of the time-seeded rand() the
// reference uses for tryRadius query rows (cu/tryRadius.cu:121,129).
inline std::vector<uint32_t> query_rows(uint32_t n, uint32_t R)
{
    uint64_t x = 0x5eedULL;
    if (const char *s = std::getenv("CASPHY_SEED")) x = std::strtoull(s, nullptr, 10);
    std::vector<uint32_t> q(R);
    for (uint32_t i = 0; i < R; ++i) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        q[i] = (uint32_t)(z % n);
    }
    return q;
}

struct Timer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double s() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

inline int device_index()
{
    if (const char *s = std::getenv("CASPHY_DEVICE")) return std::atoi(s);
    return 0;
}

}  // namespace tool
