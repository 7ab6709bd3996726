// context.hpp -- per-thread device context and shared host helpers.
#pragma once

#include <vector>

#include "common.hpp"
#include "tiles.hpp"

namespace casphy {

// Per-thread stream on the thread's selected device (default priority).
hipStream_t thread_stream();
int thread_device();

// Bit planes of integer site weights for the AND metric.
struct Planes {
    u32 nplanes = 0;             // 0 = unit weights
    std::vector<u32> mask;       // nplanes x ldw
    u32 shift[MAX_PLANES] = {};
    u32 tail = 0;
};

// Validates .w doubles (non-negative integers, sum < 2^24) and builds the
// planes for `nsite` sites laid out in rows of `ldw` words.  NULL = unit.
Planes make_planes(const double *w, u32 nsite, u32 ldw);
// Same for per-bit weights given as a vector (one-hot indel sites).
Planes make_planes_bits(const std::vector<double> &w, u32 ldw);

// Copies planes into TileArgs (uploading the masks into `buf`).
void apply_planes(const Planes &p, TileArgs &a, DevBuf &buf, hipStream_t s);

inline u32 round_up(u32 x, u32 m) { return (x + m - 1) / m * m; }

// H2D / D2H helpers on the thread stream (synchronous w.r.t. the host buffer).
void h2d(void *dst, const void *src, size_t bytes, hipStream_t s);
void d2h(void *dst, const void *src, size_t bytes, hipStream_t s);

}  // namespace casphy
