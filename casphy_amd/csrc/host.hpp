// host.hpp -- definitions shared by the library's host code and its device
// code: integer types, the error exception and the argument checks.  No HIP
// header: host-only sources (tree planning, Newick, the request loop's
// bookkeeping) also build with a plain C++ compiler.
//
// Error convention of the C ABI: every entry point returns 0 on success or a
// negative CASPHY_E* code, and casphy_last_error() returns a thread-local
// message.  No C++ exception crosses the ABI.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/casphy_hip.h"

namespace casphy {

using u32 = uint32_t;
using u64 = uint64_t;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define CASPHY_REQUIRE(cond, code, msg)                                                     \
    do {                                                                                    \
        if (!(cond)) throw ::casphy::Error((code), (msg));                                  \
    } while (0)

inline u32 ceil_div(u64 a, u64 b) { return (u32)((a + b - 1) / b); }

}  // namespace casphy
