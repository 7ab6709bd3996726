// common.hpp -- shared definitions of the CasPhy MI355X library (HIP side;
// the integer types, Error and CASPHY_REQUIRE are in host.hpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include "host.hpp"

namespace casphy {

void set_last_error(const std::string &msg);

#define CASPHY_HIP(call)                                                                    \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::casphy::Error(CASPHY_EHIP, std::string(#call) + ": " + hipGetErrorString(e_) + \
                                                   " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
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

// Per-thread stream on the thread's selected device (defined in capi.hip).
hipStream_t thread_stream();

// Inclusive prefix sum over the 64 lanes of a wave with DPP moves only (no LDS
// crossbar round trips): scans within each 16-lane row (row_shr 1, 2, 4, 8,
// zero filled), then row 0's total into row 1 and row 2's into row 3
// (row_bcast:15), then rows 0-1's total into rows 2-3 (row_bcast:31).  Every
// lane of the wave must be active.  Lane l ends with the sum over lanes 0 .. l,
// lane 63 with the wave's total.
__device__ __forceinline__ u32 wave_inclusive_scan(u32 v)
{
    v += __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xF, 0xF, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xA, 0xF, false);  // row_bcast:15 -> rows 1, 3
    v += __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xC, 0xF, false);  // row_bcast:31 -> rows 2, 3
    return v;
}

// Device buffer owned by the library (RAII).  Allocation and release are
// stream-ordered (hipMallocAsync / hipFreeAsync on the allocating thread's
// stream, from a pool that keeps freed memory), so concurrent callers on
// different streams never wait on a device-wide free.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipStream_t st = nullptr;
    DevBuf() = default;
    explicit DevBuf(size_t b) { alloc(b); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes), st(o.st) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p; bytes = o.bytes; st = o.st;
            o.p = nullptr; o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() noexcept
    {
        if (p) (void)hipFreeAsync(p, st);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t b)
    {
        release();
        if (b) {
            st = thread_stream();
            hipError_t e = hipMallocAsync(&p, b, st);
            if (e != hipSuccess) {
                p = nullptr;
                (void)hipGetLastError();
                throw Error(CASPHY_ENOMEM, "hipMallocAsync(" + std::to_string(b) + " B) failed: " +
                                                hipGetErrorString(e));
            }
            bytes = b;
        }
    }
    // grow-only reuse
    void reserve(size_t b) { if (b > bytes) alloc(b); }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

}  // namespace casphy
