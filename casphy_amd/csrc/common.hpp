// common.hpp -- shared definitions of the CasPhy MI355X library.
//
// Error convention of the C ABI: every entry point returns 0 on success or a
// negative CASPHY_E* code, and casphy_last_error() returns a thread-local
// message.  No C++ exception crosses the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
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

void set_last_error(const std::string &msg);

#define CASPHY_HIP(call)                                                                    \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::casphy::Error(CASPHY_EHIP, std::string(#call) + ": " + hipGetErrorString(e_) + \
                                                   " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

#define CASPHY_REQUIRE(cond, code, msg)                                                     \
    do {                                                                                    \
        if (!(cond)) throw ::casphy::Error((code), (msg));                                  \
    } while (0)

// Runs `body`, mapping exceptions to status codes and the thread-local message.
template <class F>
int guarded(F &&body)
{
    try {
        body();
        return CASPHY_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return CASPHY_ENOMEM;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return CASPHY_EINTERNAL;
    }
}

inline u32 ceil_div(u64 a, u64 b) { return (u32)((a + b - 1) / b); }

// Device buffer owned by the library (RAII).
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t b) { alloc(b); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) {
            if (p) (void)hipFree(p);
            p = o.p; bytes = o.bytes;
            o.p = nullptr; o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { if (p) (void)hipFree(p); }
    void alloc(size_t b)
    {
        if (p) { (void)hipFree(p); p = nullptr; }
        bytes = b;
        if (b) {
            hipError_t e = hipMalloc(&p, b);
            if (e != hipSuccess) {
                p = nullptr;
                throw Error(CASPHY_ENOMEM, "hipMalloc(" + std::to_string(b) + " B) failed: " +
                                                hipGetErrorString(e));
            }
        }
    }
    // grow-only reuse
    void reserve(size_t b) { if (b > bytes) alloc(b); }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

}  // namespace casphy
